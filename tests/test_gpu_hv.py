"""GPU parity: sector construction and H.v against the oracle.

Sector lists (list_s ranks, list_r norms) must be bit-identical to the
oracle's restatement of allocate_lists_omp_SQ (state_lists.f90:225-292); one
H.v in FP64 must agree with ham_to_vec_wave_vector (ham2vec_v3.f90:37-72)
within 1e-12 relative (max-norm).  The FP32 path computes (H - e_zz) v in
complex float; it is checked against the oracle's FP64 result at FP32
tolerance (2e-6 relative to max |(H - e_zz) v|)."""
import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _q3():
    from ed_amd import qs3 as q3
    return q3


def chain(L, fam):
    return orc.chain_bonds(L, fam)


def dimer(L, j1, j2, jz1, jz2):
    """Nearest-neighbour chain with alternating bonds (two bond types)."""
    return [(i, i % L + 1, j1 if i % 2 else j2, jz1 if i % 2 else jz2) for i in range(1, L + 1)]


CASES = [
    # name, NOS, LX, LY, LZ, nod, (KX, KY, KZ), bonds
    ("chain12_n4_k0", 12, 12, 1, 1, 4, (0, 0, 0), chain(12, [(1, 1.0, 1.0)])),
    ("chain12_n4_k3_j1j2", 12, 12, 1, 1, 4, (3, 0, 0), chain(12, [(1, 1.0, 0.7), (2, 0.5, 0.5)])),
    ("chain10_n5_k5", 10, 10, 1, 1, 5, (5, 0, 0), chain(10, [(1, 1.0, 1.0)])),
    ("chain16_n8_k0", 16, 16, 1, 1, 8, (0, 0, 0), chain(16, [(1, 1.0, 1.0)])),
    ("chain16_n8_k4", 16, 16, 1, 1, 8, (4, 0, 0), chain(16, [(1, 1.0, 0.8)])),
    ("chain20_n1_k3", 20, 20, 1, 1, 1, (3, 0, 0), chain(20, [(1, 1.0, 1.0)])),
    ("chain24_n3_k7_nnn", 24, 24, 1, 1, 3, (7, 0, 0), chain(24, [(1, 1.0, 1.0), (2, 0.5, 0.5)])),
    ("chain36_n6_k0_C1", 36, 36, 1, 1, 6, (0, 0, 0), chain(36, [(1, 1.0, 1.0)])),
    ("chain14_n7_k0", 14, 14, 1, 1, 7, (0, 0, 0), chain(14, [(1, 1.0, 1.0)])),
    ("chain14_n7_k7", 14, 14, 1, 1, 7, (7, 0, 0), chain(14, [(1, 1.0, 0.6)])),
    ("chain15_n5_k3_dimer", 15, 15, 1, 1, 5, (3, 0, 0), dimer(15, 1.0, 0.4, 0.8, 0.3)),
    ("chain16_n8_k2_dimer", 16, 16, 1, 1, 8, (2, 0, 0), dimer(16, 1.0, 0.5, 1.0, 0.25)),
    ("chain8_n2_k4", 8, 8, 1, 1, 2, (4, 0, 0), chain(8, [(1, 1.0, 1.0)])),
    ("chain3_n2_k1", 3, 3, 1, 1, 2, (1, 0, 0), chain(3, [(1, 1.0, 0.5)])),
    ("chain40_n3_k5", 40, 40, 1, 1, 3, (5, 0, 0), chain(40, [(1, 0.7, 1.3)])),
    ("torus6x6_n3_k00", 36, 6, 6, 1, 3, (0, 0, 0), orc.square_bonds(6, 6, 1.0)),
    ("torus6x6_n3_k12", 36, 6, 6, 1, 3, (1, 2, 0), orc.square_bonds(6, 6, 1.0)),
    ("torus4x4_n4_k10", 16, 4, 4, 1, 4, (1, 0, 0), orc.square_bonds(4, 4, 1.0)),
    ("cube3x3x3_n2_k111", 27, 3, 3, 3, 2, (1, 1, 1), [(1, 2, 1.0, 0.5), (1, 4, 1.0, 0.5), (1, 10, 0.5, 1.0), (5, 14, 1.0, 1.0)]),
]


def _build(case):
    name, NOS, LX, LY, LZ, nod, K, bonds = case
    q3 = _q3()
    om = orc.Model(NOS, LX, LY, LZ, bonds)
    rk = om.rk_from_K(*K)
    osec = om.build_sector(nod, rk)
    gm = q3.Model(NOS, LX, LY, LZ, bonds)
    gsec = q3.Sector.build(gm, nod, rk)
    return om, osec, gm, gsec


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_sector_lists_bit_exact(case):
    print(f"[gpu] sector {case[0]}", flush=True)
    om, osec, gm, gsec = _build(case)
    assert gsec.dim == osec.dim
    ls, lr = gsec.lists()
    assert np.array_equal(ls, osec.list_s)
    assert np.array_equal(lr, osec.list_r), np.max(np.abs(lr - osec.list_r))


# 0: fastest (nearest-neighbour chain where it applies), 1: batched-load chain,
# 2: per-hop chain, 3: general translation scan
KERNELS = [0, 1, 2, 3]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hv_fp64_parity(case, kernel):
    print(f"[gpu] hv fp64 {case[0]} kernel={kernel}", flush=True)
    q3 = _q3()
    om, osec, gm, gsec = _build(case)
    rng = np.random.default_rng(7)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref = om.hv(osec, v)
    q3.set_option("hv_kernel", kernel)
    try:
        got = gsec.ham_to_vec_wave_vector(v)
    finally:
        q3.set_option("hv_kernel", 0)
    err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300)
    assert err < 1e-12, err
    # nnz matches the oracle's count of in-sector hop targets
    assert gsec.nnz() == om.count_hops(osec)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hv_fp32_parity(case, kernel):
    print(f"[gpu] hv fp32 {case[0]} kernel={kernel}", flush=True)
    q3 = _q3()
    om, osec, gm, gsec = _build(case)
    rng = np.random.default_rng(11)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref = om.hv(osec, v) - om.fm_energy() * v
    dv = torch.from_numpy(v.astype(np.complex64)).cuda()
    out = torch.empty_like(dv)
    q3.set_option("hv_kernel", kernel)
    try:
        gsec.hv_dev(out.data_ptr(), dv.data_ptr(), q3.FP32)
        torch.cuda.synchronize()
    finally:
        q3.set_option("hv_kernel", 0)
    got = out.cpu().numpy().astype(np.complex128)
    err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300)
    assert err < 2e-6, err


def test_hv_from_reference_lists():
    """Drop-in use: a sector given as (list_s, list_r, explist) produced
    elsewhere (here the oracle) gives the same H.v."""
    print("[gpu] hv from lists", flush=True)
    q3 = _q3()
    for case in (CASES[1], CASES[9]):
        name, NOS, LX, LY, LZ, nod, K, bonds = case
        om, osec, gm, _ = _build(case)
        sec = q3.Sector.from_lists(gm, nod, osec.list_s, osec.list_r, osec.explist)
        rng = np.random.default_rng(3)
        v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
        ref = om.hv(osec, v)
        got = sec.ham_to_vec_wave_vector(v)
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("case", [c for c in CASES if c[0] in ("chain12_n4_k3_j1j2", "chain16_n8_k4", "chain10_n5_k5",
                                                             "torus6x6_n3_k12")], ids=lambda c: c[0])
def test_hv_autotuned_parity(case):
    """qs3x_hv_autotune keeps the fastest launch configuration per sector and
    precision; H.v through that configuration still equals the oracle's (FP64
    1e-12, FP32 2e-6), and clearing the tuning restores the default."""
    print(f"[gpu] hv autotuned {case[0]}", flush=True)
    q3 = _q3()
    om, osec, gm, gsec = _build(case)
    rng = np.random.default_rng(13)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref = om.hv(osec, v)
    best64, ms64, res64 = gsec.hv_autotune(q3.FP64, n_rep=2)
    assert len(res64) >= 2 and ms64 > 0 and best64["xcd"] in (1, 8) and best64["grid_quarters"] in (1, 2, 3, 4)
    assert ms64 <= min(r[11] for r in res64 if r[11] > 0) / 0.98
    # stage 1 timed each family once, stage 2 the winner's parameters
    fams = [int(r[0]) for r in res64]
    assert best64["family"] in fams and len(set(fams)) == len([f for f in dict.fromkeys(fams)])
    got = gsec.ham_to_vec_wave_vector(v)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12
    best32, ms32, _ = gsec.hv_autotune(q3.FP32, n_rep=2)
    ref32 = ref - om.fm_energy() * v
    dv = torch.from_numpy(v.astype(np.complex64)).cuda()
    out = torch.empty_like(dv)
    gsec.hv_dev(out.data_ptr(), dv.data_ptr(), q3.FP32)
    torch.cuda.synchronize()
    got32 = out.cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got32 - ref32)) / np.max(np.abs(ref32)) < 2e-6
    gsec.hv_tune_clear()
    got = gsec.ham_to_vec_wave_vector(v)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("prec_name", ["fp32", "fp64"])
def test_hv_autotune_stencil_sector(prec_name):
    """On a stencil sector (nearest-neighbour chain, N_down = 4) the tuner's
    second stage searches the stencil's kernel form, tile height, task slabs
    and windows in flight (or the nearest-neighbour kernel's rows per lane, chunk,
    schedule and grid); the chosen configuration is the one later H.v use,
    and it still equals the oracle's H.v (FP64 1e-12, FP32 2e-6)."""
    print(f"[gpu] hv autotune stencil sector {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    L, K = 160, 7
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.9)])
    om = orc.Model(L, L, 1, 1, bonds)
    rk = om.rk_from_K(K)
    gm = q3.Model(L, L, 1, 1, bonds)
    s = q3.Sector.build(gm, 4, rk)
    best, ms, res = s.hv_autotune(prec, n_rep=3)
    fams = [int(r[0]) for r in res]
    assert 5 in fams and 0 in fams  # the stencil and the nearest-neighbour kernel were both timed
    n1 = len(set(fams))  # stage 1: one row per family
    stage2 = res[n1:]
    assert len(stage2) >= 2 and all(int(r[0]) == best["family"] for r in stage2)
    if best["family"] == 5:
        # kernel forms (default and slab window), heights, windows in flight
        fdef = 1 if prec == q3.FP64 else 2
        assert {int(r[10]) for r in stage2} == {fdef, 0}
        assert {6, 8, 10} <= {int(r[3]) for r in stage2} and {2, 3} <= {int(r[5]) for r in stage2}
    rng = np.random.default_rng(K)
    v = rng.uniform(-1, 1, s.dim) + 1j * rng.uniform(-1, 1, s.dim)
    osec = om.build_sector(4, rk)
    ref = om.hv(osec, v)
    if prec == q3.FP64:
        got = s.ham_to_vec_wave_vector(v)
        assert s.hv_last_launch()[0] == best["family"]
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    else:
        ref32 = ref - om.fm_energy() * v
        dv = torch.from_numpy(v.astype(np.complex64)).cuda()
        out = torch.empty_like(dv)
        s.hv_dev(out.data_ptr(), dv.data_ptr(), q3.FP32)
        torch.cuda.synchronize()
        assert s.hv_last_launch()[0] == best["family"]
        got32 = out.cpu().numpy().astype(np.complex128)
        assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32))
    if best["family"] == 5 and best["lines"] > 0:
        # the plan of that launch is the tuned tile
        info = s.hv_stencil_info(prec)
        assert info["tile_lines"] == best["lines"] and info["slabs"] == best["slabs"]
        assert info["form"] == best["form"]
    s.hv_tune_clear()


@pytest.mark.parametrize("prec_name", ["fp32", "fp64"])
def test_hv_tile_schedules_bit_identical(prec_name):
    """Every tile schedule of the nearest-neighbour kernels (round robin, one
    range per XCD, interleaved chunks of 2^c tiles) writes every row exactly
    once with the same value: each row is computed by one lane in a fixed
    order, so the outputs are bit-identical.  The output is filled with NaN
    first, so a row that no block visits would show.  Chain L=120, N_down=4
    (about 70,000 rows: many tiles, a partial last chunk, and chunks larger
    than the sector), k=0 and k=7."""
    print(f"[gpu] hv tile schedules {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP32 if prec_name == "fp32" else q3.FP64
    dt = torch.complex64 if prec_name == "fp32" else torch.complex128
    L = 120
    gm = q3.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.9)]))
    try:
        for K in (0, 7):
            s = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
            g = torch.Generator(device="cuda")
            g.manual_seed(K + 1)
            x = torch.randn(s.dim, dtype=dt, device="cuda", generator=g)
            outs = []
            for xcd, chunk in [(0, -1), (1, 0), (8, 0), (8, 1), (8, 3), (8, 9), (3, 2), (64, 0)]:
                q3.set_option("hv_xcd", xcd)
                q3.set_option("hv_chunk", chunk)
                y = torch.full_like(x, float("nan"))
                s.hv_dev(y.data_ptr(), x.data_ptr(), prec)
                torch.cuda.synchronize()
                outs.append(y.cpu().numpy())
            assert not np.isnan(outs[0]).any()
            for o in outs[1:]:
                assert np.array_equal(o, outs[0])
    finally:
        q3.set_option("hv_xcd", 0)
        q3.set_option("hv_chunk", -1)


@pytest.mark.parametrize("case", [c for c in CASES if c[0] in ("chain12_n4_k3_j1j2", "chain10_n5_k5", "chain40_n3_k5",
                                                             "chain15_n5_k3_dimer", "chain36_n6_k0_C1")],
                         ids=lambda c: c[0])
def test_hv_fp32_unaligned_vectors(case):
    """FP32 H.v on vectors that are 8-byte but not 16-byte aligned (as the
    odd slots of a thick-restart basis are when D is odd) runs the
    one-row-per-lane kernel instead of the two-rows-per-lane one; both equal
    the oracle's (H - e_zz) v within 2e-6 relative."""
    print(f"[gpu] hv fp32 unaligned {case[0]}", flush=True)
    q3 = _q3()
    om, osec, gm, gsec = _build(case)
    rng = np.random.default_rng(17)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref32 = om.hv(osec, v) - om.fm_energy() * v
    D = gsec.dim
    for off in (0, 1):
        xb = torch.zeros(D + 1, dtype=torch.complex64, device="cuda")
        yb = torch.full((D + 1,), float("nan"), dtype=torch.complex64, device="cuda")
        x, y = xb[off:off + D], yb[off:off + D]
        x.copy_(torch.from_numpy(v.astype(np.complex64)).cuda())
        assert (x.data_ptr() % 16 == 0) == (off == 0)
        gsec.hv_dev(y.data_ptr(), x.data_ptr(), q3.FP32)
        torch.cuda.synchronize()
        got = y.cpu().numpy().astype(np.complex128)
        assert np.max(np.abs(got - ref32)) <= 2e-6 * np.max(np.abs(ref32)), off


LARGE_CHAINS = [
    # name, L, N_down, K, bonds: sectors of 6,567 to 119,133 rows, where a
    # wavefront's index words span several bitmap words, the per-XCD chunks
    # are many, and runs of consecutive rows are long
    ("chain128_n4_k7_dimer", 128, 4, 7, dimer(128, 1.0, 0.6, 0.9, 0.4)),
    ("chain64_n5_k0", 64, 5, 0, chain(64, [(1, 1.0, 1.0)])),
    ("chain100_n4_k50", 100, 4, 50, chain(100, [(1, 1.0, 0.7)])),
    ("chain200_n3_k1", 200, 3, 1, chain(200, [(1, 1.0, 1.0)])),
]


@pytest.mark.parametrize("case", LARGE_CHAINS, ids=[c[0] for c in LARGE_CHAINS])
def test_hv_large_chain_parity(case):
    """The nearest-neighbour kernels on larger chains against the oracle:
    FP64 within 1e-12 and FP32 (the two-rows-per-lane kernel) within 2e-6
    relative, and the sector lists bit-identical."""
    print(f"[gpu] hv large chain {case[0]}", flush=True)
    q3 = _q3()
    name, L, N, K, bonds = case
    om = orc.Model(L, L, 1, 1, bonds)
    rk = om.rk_from_K(K, 0, 0)
    osec = om.build_sector(N, rk)
    gm = q3.Model(L, L, 1, 1, bonds)
    gsec = q3.Sector.build(gm, N, rk)
    assert gsec.dim == osec.dim
    ls, lr = gsec.lists()
    assert np.array_equal(ls, osec.list_s) and np.array_equal(lr, osec.list_r)
    rng = np.random.default_rng(23)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref = om.hv(osec, v)
    got = gsec.ham_to_vec_wave_vector(v)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    ref32 = ref - om.fm_energy() * v
    dv = torch.from_numpy(v.astype(np.complex64)).cuda()
    out = torch.full_like(dv, float("nan"))
    gsec.hv_dev(out.data_ptr(), dv.data_ptr(), q3.FP32)
    torch.cuda.synchronize()
    got32 = out.cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32))


@pytest.mark.parametrize("L,K", [(8000, 0), (8000, 3), (20000, 1)])
def test_hv_nn_long_chain_matches_general(L, K):
    """Long chains (N_down = 2): the nearest-neighbour kernel runs on its own
    LDS tables (48 KB at L = 8000; about 120 KB at L = 20000, beyond the
    default 64 KB) where the general chain kernels' tables would not fit.
    Its H.v equals the general translation-scan kernel's (option hv_kernel
    3, pinned against the oracle on the smaller lattices): FP64 within 1e-12,
    FP32 within 2e-6 relative.  The oracle's own sector build is too slow at
    these lengths."""
    print(f"[gpu] hv long chain L={L} k={K}", flush=True)
    q3 = _q3()
    gm = q3.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.8)]))
    s = q3.Sector.build(gm, 2, (q3.wavevector(K, L), 0.0, 0.0))
    rng = np.random.default_rng(29)
    v = rng.uniform(-1, 1, s.dim) + 1j * rng.uniform(-1, 1, s.dim)
    out = {}
    try:
        for kern in (0, 3):
            q3.set_option("hv_kernel", kern)
            out[(kern, 64)] = s.ham_to_vec_wave_vector(v)
            fam64, lds64 = s.hv_last_launch()
            dv = torch.from_numpy(v.astype(np.complex64)).cuda()
            y = torch.full_like(dv, float("nan"))
            s.hv_dev(y.data_ptr(), dv.data_ptr(), q3.FP32)
            torch.cuda.synchronize()
            out[(kern, 32)] = y.cpu().numpy().astype(np.complex128)
            # the kernel asked for is the one that ran (no silent fallback)
            assert s.hv_last_launch()[0] == fam64 == kern, (kern, fam64, s.hv_last_launch())
            if kern == 0 and L == 20000:
                assert lds64 > 64 * 1024
    finally:
        q3.set_option("hv_kernel", 0)
    ref = out[(3, 64)]
    assert np.max(np.abs(out[(0, 64)] - ref)) <= 1e-12 * np.max(np.abs(ref))
    ref32 = ref - (gm.bonds[:, 3].sum() / 4.0) * v
    for kern in (0, 3):
        assert np.max(np.abs(out[(kern, 32)] - ref32)) <= 2e-6 * np.max(np.abs(ref32)), kern


@pytest.mark.parametrize("K", [0, 2])
def test_hv_long_j1j2_chain_matches_general(K):
    """A long J1-J2 chain (L = 4000, N_down = 2; about 96 KB of LDS tables):
    the batched-load and per-hop chain kernels (options hv_kernel 1 and 2)
    equal the general translation-scan kernel (hv_kernel 3) within 1e-12 in
    FP64 and 2e-6 in FP32."""
    print(f"[gpu] hv long j1j2 chain k={K}", flush=True)
    q3 = _q3()
    L = 4000
    gm = q3.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.8), (2, 0.4, 0.3)]))
    s = q3.Sector.build(gm, 2, (q3.wavevector(K, L), 0.0, 0.0))
    rng = np.random.default_rng(31)
    v = rng.uniform(-1, 1, s.dim) + 1j * rng.uniform(-1, 1, s.dim)
    out = {}
    try:
        for kern in (1, 2, 3):
            q3.set_option("hv_kernel", kern)
            out[(kern, 64)] = s.ham_to_vec_wave_vector(v)
            fam64, lds64 = s.hv_last_launch()
            dv = torch.from_numpy(v.astype(np.complex64)).cuda()
            y = torch.full_like(dv, float("nan"))
            s.hv_dev(y.data_ptr(), dv.data_ptr(), q3.FP32)
            torch.cuda.synchronize()
            out[(kern, 32)] = y.cpu().numpy().astype(np.complex128)
            # the kernel asked for ran, with its tables beyond 64 KB of LDS
            fam32, lds32 = s.hv_last_launch()
            assert fam64 == fam32 == kern, (kern, fam64, fam32)
            if kern in (1, 2):
                assert lds64 > 64 * 1024 and lds32 > 64 * 1024, (kern, lds64, lds32)
    finally:
        q3.set_option("hv_kernel", 0)
    ref = out[(3, 64)]
    ref32 = ref - (gm.bonds[:, 3].sum() / 4.0) * v
    for kern in (1, 2):
        assert np.max(np.abs(out[(kern, 64)] - ref)) <= 1e-12 * np.max(np.abs(ref)), kern
        assert np.max(np.abs(out[(kern, 32)] - ref32)) <= 2e-6 * np.max(np.abs(ref32)), kern


STENCIL_CASES = [
    # L, K, Jz, hv_stencil_lines, hv_stencil_slabs, hv_stencil_depth
    (40, 0, 0.8, 0, 16, 0),
    (40, 11, 0.8, 6, 5, 3),
    (64, 3, 1.0, 2, 1, 0),
    (64, 0, 1.0, 8, 16, 3),
    (128, 0, 1.0, 0, 0, 0),
    (128, 17, 0.6, 4, 7, 3),
]


@pytest.mark.parametrize("L,K,jz,B,slabs,depth", STENCIL_CASES)
def test_hv_stencil_parity(L, K, jz, B, slabs, depth):
    """Gap-lattice stencil H.v (nearest-neighbour chains, uniform bonds,
    N_down = 4): FP64 within 1e-12 of the oracle's ham_to_vec_wave_vector
    (ham2vec_v3.f90:37-72), FP32 (H - e_zz) v within 2e-6, for several band
    widths, task lengths and prefetch depths, with 16-byte aligned and
    unaligned FP32 vectors; the stencil kernel is the one that ran."""
    print(f"[gpu] hv stencil L={L} K={K} B={B} slabs={slabs}", flush=True)
    q3 = _q3()
    bonds = orc.chain_bonds(L, [(1, 1.0, jz)])
    om = orc.Model(L, L, 1, 1, bonds)
    rk = om.rk_from_K(K)
    osec = om.build_sector(4, rk)
    gm = q3.Model(L, L, 1, 1, bonds)
    q3.set_option("hv_stencil_lines", B)
    q3.set_option("hv_stencil_slabs", slabs)
    q3.set_option("hv_stencil_depth", depth)
    try:
        s = q3.Sector.build(gm, 4, rk)
        rng = np.random.default_rng(L + K)
        v = rng.uniform(-1, 1, s.dim) + 1j * rng.uniform(-1, 1, s.dim)
        ref = om.hv(osec, v)
        got = s.ham_to_vec_wave_vector(v)
        assert s.hv_last_launch()[0] == 5
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
        ref32 = ref - om.fm_energy() * v
        for shift in (0, 1):  # element offset 1: v1 and v0 only 8-byte aligned
            buf = torch.from_numpy(np.concatenate([np.zeros(shift), v]).astype(np.complex64)).cuda()
            dv = buf[shift:]
            ybuf = torch.full((s.dim + shift,), float("nan"), dtype=torch.complex64, device="cuda")
            y = ybuf[shift:]
            s.hv_dev(y.data_ptr(), dv.data_ptr(), q3.FP32)
            torch.cuda.synchronize()
            assert s.hv_last_launch()[0] == 5
            got32 = y.cpu().numpy().astype(np.complex128)
            assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32)), shift
    finally:
        q3.set_option("hv_stencil_lines", 0)
        q3.set_option("hv_stencil_slabs", 0)
        q3.set_option("hv_stencil_depth", 0)


def test_hv_stencil_matches_nn_kernel_c3():
    """At L = 1000, N_down = 4 (41,417,250 rows) the stencil H.v and the
    nearest-neighbour chain kernel give the same result (FP64 1e-13 of the
    largest element, FP32 2e-6), k = 0 and k = 2 pi 7 / 1000."""
    q3 = _q3()
    L = 1000
    gm = q3.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    for K in (0, 7):
        print(f"[gpu] hv stencil vs nn kernel L=1000 K={K}", flush=True)
        s = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
        g = torch.Generator(device="cuda")
        g.manual_seed(5 + K)
        for dt, tol in ((torch.complex128, 1e-13), (torch.complex64, 2e-6)):
            prec = q3.FP64 if dt == torch.complex128 else q3.FP32
            x = torch.randn(s.dim, dtype=dt, device="cuda", generator=g)
            y1 = torch.empty_like(x)
            y0 = torch.empty_like(x)
            s.hv_dev(y1.data_ptr(), x.data_ptr(), prec)
            assert s.hv_last_launch()[0] == 5
            q3.set_option("hv_stencil", 0)
            try:
                s.hv_dev(y0.data_ptr(), x.data_ptr(), prec)
                assert s.hv_last_launch()[0] == 0
            finally:
                q3.set_option("hv_stencil", 1)
            torch.cuda.synchronize()
            err = ((y1 - y0).abs().max() / y0.abs().max()).item()
            assert err <= tol, (K, dt, err)
            del x, y0, y1
        del s


def test_sector_build_refuses_incommensurate_wave_vector():
    """rk L / (2 pi) must be an integer in every direction (the target sector
    of the DSF is built at k + q, get_expectation_values.f90:430-447)."""
    print("[gpu] incommensurate wave vector", flush=True)
    q3 = _q3()
    gm = q3.Model(12, 12, 1, 1, orc.chain_bonds(12, [(1, 1.0, 1.0)]))
    with pytest.raises(q3.QS3XError, match="not a multiple of 2 pi"):
        q3.Sector.build(gm, 3, (2 * np.pi * 2.5 / 12, 0.0, 0.0))
    gt = q3.Model(16, 4, 4, 1, orc.square_bonds(4, 4, 1.0))
    with pytest.raises(q3.QS3XError, match="rky"):
        q3.Sector.build(gt, 2, (0.0, 0.3, 0.0))
    # commensurate vectors still build, including 2 pi K / L with K >= L
    s = q3.Sector.build(gm, 3, (q3.wavevector(17, 12), 0.0, 0.0))
    assert s.dim == q3.Sector.build(gm, 3, (q3.wavevector(5, 12), 0.0, 0.0)).dim
