"""GPU tests at the full sizes of the three large workloads.

C3: periodic NN Heisenberg chain L = 1000, N_down = 4, k = 0 (41,417,250 rows):
    the sector (allocate_lists_omp_SQ, state_lists.f90:225-292) checked on
    windows against the oracle's checkstate (state_lists.f90:66-112); one H.v
    (ham_to_vec_wave_vector, ham2vec_v3.f90:37-72) checked on windows of rows
    against the oracle's row-restricted restatement (FP64 1e-12, FP32 2e-6);
    plain Lanczos (lanczos.f90:50-152) within 1e-9 of the Bethe-ansatz energy
    242.000099091345514, in FP64 and in the FP32-mixed recurrence.
C4: periodic XXZ chain L = 40, Jz = 0.8, every (N_down = 1..7, k) sector: all
    280 dimensions against the character formula, and the batched group solve
    against sector-by-sector solves in FP64 (1e-9).
C5: the lowest k = 0 state of N_down = 3 at L = 1000, S^-_q into the
    N_down = 4, k = q sectors for the 64 wave numbers K_m = round(1000 m / 64):
    the 64-q block recurrence (calcu_DSF_wavevector,
    get_expectation_values.f90:379-605) against the per-q path at four q, one
    of them (K = 16, q = 2 pi 16 / 1000) snapped from 2 pi / 64.
"""
import math

import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

E0_C3 = 242.000099091345514  # Bethe ansatz, L = 1000, M = 4
D_C3 = 41417250


def _q3():
    from ed_amd import qs3 as q3
    return q3


def char_dim(L, N, K):
    """Dimension of the (N, k = 2 pi K / L) sector of a ring: (1/L) sum_r
    e^{-2 pi i K r / L} Fix(T^r), Fix(T^r) = C(g, N g / L) when L/g divides N
    (g = gcd(r, L)), else 0."""
    s = 0.0
    for r in range(L):
        g = math.gcd(r, L)
        if (N * g) % L == 0:
            s += math.cos(2 * math.pi * K * r / L) * math.comb(g, N * g // L)
    return int(round(s / L))


@pytest.fixture(scope="module")
def c3():
    print("[gpu] C3 sector build L=1000 N=4 k=0", flush=True)
    q3 = _q3()
    L = 1000
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    gm = q3.Model(L, L, 1, 1, bonds)
    sec = q3.Sector.build(gm, 4, (0.0, 0.0, 0.0))
    ls, lr = sec.lists()
    om = orc.Model(L, L, 1, 1, bonds)
    osec = orc.Sector(4, (0.0, 0.0, 0.0), ls, lr, orc.explist(L, 1, 1, 0.0, 0.0, 0.0))
    yield q3, gm, sec, om, osec
    del sec


def test_c3_dimension_and_windows_of_the_list(c3):
    q3, gm, sec, om, osec = c3
    assert sec.dim == D_C3 == char_dim(1000, 4, 0)
    ls, lr = osec.list_s, osec.list_r
    assert np.all(np.diff(ls) > 0)
    # every rank between the first and last row of a window is a representative
    # exactly when it is listed, with list_r = sqrt(ro) bit for bit
    # (windows of up to 40 rows, cut where they would span over 200,000 ranks)
    for i0 in (0, 777_777, D_C3 // 2, D_C3 - 40):
        i1 = min(i0 + 40, int(np.searchsorted(ls, ls[i0] + 200_000)))
        assert i1 - i0 >= 2
        listed = {int(ls[i]): float(lr[i]) for i in range(i0, i1)}
        for s in range(int(ls[i0]), int(ls[i1 - 1]) + 1):
            ro, _ = om.checkstate(s, 4, (0.0, 0.0, 0.0))
            if s in listed:
                assert ro > 1e-15 and math.sqrt(ro) == listed[s], s
            else:
                assert ro <= 1e-15, s


def test_c3_hv_windows_fp64_and_fp32(c3):
    q3, gm, sec, om, osec = c3
    print("[gpu] C3 H.v windows", flush=True)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    x64 = torch.randn(D_C3, dtype=torch.complex128, device=dev, generator=g)
    y64 = torch.empty_like(x64)
    sec.hv_dev(y64.data_ptr(), x64.data_ptr(), q3.FP64)
    x32 = x64.to(torch.complex64)
    y32 = torch.empty_like(x32)
    sec.hv_dev(y32.data_ptr(), x32.data_ptr(), q3.FP32)
    torch.cuda.synchronize()
    v = x64.cpu().numpy()
    v32 = x32.cpu().numpy().astype(np.complex128)
    ezz = om.fm_energy()
    assert ezz == 250.0
    W = 4096
    for s in (0, 12_345_678, D_C3 // 2, D_C3 - W):
        ref = om.hv(osec, v, rows=(s, s + W))[s:s + W]
        got = y64[s:s + W].cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), s
        ref32 = om.hv(osec, v32, rows=(s, s + W))[s:s + W] - ezz * v32[s:s + W]
        got32 = y32[s:s + W].cpu().numpy().astype(np.complex128)
        assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32)), s


def test_c3_plain_lanczos_reaches_bethe_energy(c3):
    q3, gm, sec, om, osec = c3
    print("[gpu] C3 plain Lanczos FP64 / FP32-mixed", flush=True)
    E64, _, it64 = q3.lanczos_routines(sec, maxitr=5000, minitr=20, itrint=5, lnc_ene_conv=1e-11, prec=q3.FP64,
                                       seed=1, want_vector=False)
    assert abs(E64 / E0_C3 - 1) < 1e-9, (E64, it64)
    E32, _, it32 = q3.lanczos_routines(sec, maxitr=5000, minitr=20, itrint=5, lnc_ene_conv=1e-11, prec=q3.FP32,
                                       seed=1, want_vector=False)
    assert abs(E32 / E0_C3 - 1) < 1e-9, (E32, it32)


def test_c4_dimensions_and_group_solve_fp64():
    print("[gpu] C4 280 sectors L=40", flush=True)
    q3 = _q3()
    L = 40
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 0.8)]))
    secs = []
    for N in range(1, 8):
        for K in range(L):
            s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
            assert s.dim == char_dim(L, N, K), (N, K)
            secs.append(s)
    assert len(secs) == 280
    assert sum(s.dim for s in secs) == 23242038
    assert all(s.dim == 466089 for s in secs[-40:])
    kw = dict(maxitr=400, minitr=20, itrint=5, lnc_ene_conv=1e-12, seed=1)
    E, its, st = q3.lanczos_group(secs, prec=q3.FP64, use_graph=True, **kw)
    E1 = np.array([q3.lanczos_routines(s, prec=q3.FP64, want_vector=False, **kw)[0] for s in secs])
    assert np.max(np.abs(E - E1) / np.maximum(np.abs(E1), 1.0)) < 1e-9


@pytest.fixture(scope="module")
def c5():
    print("[gpu] C5 ground state L=1000 N=3 k=0", flush=True)
    q3 = _q3()
    L = 1000
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 1.0)]))
    gs = q3.Sector.build(gm, 3, (0.0, 0.0, 0.0))
    assert gs.dim == 166167 == char_dim(L, 3, 0)
    E0, psi, _ = q3.lanczos_routines(gs, maxitr=3000, lnc_ene_conv=1e-13, prec=q3.FP64, seed=1)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = [int(round(1000 * m / 64)) for m in range(64)]
    return q3, gm, gs, psi, pos, Ks


def _per_q(q3, gm, gs, psi, pos, K, steps, prec):
    rk = q3.wavevector(K, 1000)
    tgt = q3.Sector.build(gm, 4, (rk, 0.0, 0.0))
    f = q3.calcu_Sq_a(2, gs, tgt, np.array([rk, 0.0, 0.0]), pos, psi)
    a, b, n = q3.dsf_recurrence(tgt, f, steps, prec=prec)
    return a, b, n, tgt.dim


SAMPLE = [1, 8, 37, 63]  # K = 16 (snapped), 125 (exact 2 pi / 8), 578, 984


def test_c5_block_fp64_matches_per_q(c5):
    q3, gm, gs, psi, pos, Ks = c5
    steps = 20
    print("[gpu] C5 64-q block FP64, 20 steps", flush=True)
    a, b, n, st = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=q3.FP64)
    assert st["dim_all_k"] >= 41417250
    for j in SAMPLE:
        print(f"[gpu] C5 per-q FP64 K={Ks[j]}", flush=True)
        ap, bp, nq, dim = _per_q(q3, gm, gs, psi, pos, Ks[j], steps, q3.FP64)
        assert dim == char_dim(1000, 4, Ks[j])
        assert abs(n[j] - nq) <= 1e-12 * nq, (j, n[j], nq)
        assert np.max(np.abs(a[j] - ap)) <= 1e-8 * max(1.0, np.max(np.abs(ap))), j
        assert np.max(np.abs(b[j] - bp)) <= 1e-8 * max(1.0, np.max(np.abs(bp))), j
    # K = 0: S^-_{q=0} is the total lowering operator, so S^-|psi0> is an
    # eigenvector (SU(2)) to the accuracy of psi0: beta_2^2 = |(H - alpha_1) f|^2
    # is the squared residual of the Lanczos ground state (about 1e-12 here)
    assert b[0][1] <= 1e-9 * n[0]


def test_c5_block_fp32_early_coefficients(c5):
    """FP32 block against the FP64 block.  S^-_q |psi0> of a dilute chain has
    nearly all its weight on a few states, so beta_n^2 falls fast and the
    FP32 rounding of each step (relative to |H f|) is amplified by 1/beta:
    the FP32 coefficients follow FP64 over the first four levels only (the
    first eight spectral moments); later ones are checked in FP64."""
    q3, gm, gs, psi, pos, Ks = c5
    steps = 20
    print("[gpu] C5 64-q block FP32, 20 steps", flush=True)
    a, b, n, _ = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=q3.FP32)
    a64, b64, n64, _ = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=q3.FP64)
    assert np.max(np.abs(n - n64) / np.abs(n64)) <= 1e-12
    for j in range(1, 64):
        ea = np.max(np.abs(a[j, :4] - a64[j, :4])) / max(1.0, np.max(np.abs(a64[j, :4])))
        eb = np.max(np.abs(b[j, :4] - b64[j, :4])) / max(1e-300, np.max(np.abs(b64[j, :4])))
        assert ea <= 1e-7 and eb <= 1e-4, (j, ea, eb, b64[j, :8].tolist())
