"""GPU parity of the column form of the gap-lattice stencil H.v.

On nearest-neighbour chains with uniform bonds and N_down = 4 the H.v of
ham2vec_v3.f90:37-72 runs as an 8-point stencil on the (n2, n3, n4) lattice
of clear rows.  Option hv_stencil_kernel selects how a block's threads take
the rows of a tile: 0 the slab-window form (rows of each slab dealt to the
threads in turn), 1 the column form (a thread keeps its window positions for
the whole sweep of a task, its own value and its neighbours in n4 in
registers), 2 the pair form (FP32: the column form with two adjacent
positions per lane and 16-byte accesses; FP64 falls back to the column
form).  Both form every sum in the same order, so their results are
compared bit for bit, and the column form against the oracle's
ham_to_vec_wave_vector (FP64 1e-12, FP32 (H - e_zz) v 2e-6), in the plain
H.v and in the fused Lanczos step (lanczos.f90:50-92)."""
import numpy as np
import pytest
import torch

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu


def _q3():
    from ed_amd import qs3 as q3
    return q3


def _opts(q3, kernel, lines=0, slabs=0, span=0):
    """span: window positions per lane of the column form (tiles 64 span - 2
    wide).  The slab-window form takes at most 8 lines; a row's sums do not
    depend on the tiling, so the results do not change with it."""
    q3.set_option("hv_stencil_kernel", kernel)
    q3.set_option("hv_stencil_lines", lines if kernel == 1 else min(lines, 8))
    q3.set_option("hv_stencil_slabs", slabs)
    q3.set_option("hv_stencil_span", span if kernel == 1 else 0)


def _reset(q3):
    _opts(q3, 0)


CASES = [
    # L, K, Jz, lines per tile (0: default), slabs per task (0: default),
    # window positions per lane (0: default)
    (20, 0, 1.0, 0, 0, 0),
    (40, 11, 0.8, 6, 5, 2),
    (64, 3, 1.0, 2, 1, 2),
    (64, 0, 1.0, 8, 16, 1),
    (128, 0, 1.0, 0, 0, 2),
    (128, 17, 0.6, 4, 7, 2),
    (203, 5, 1.0, 8, 3, 2),
    (150, 0, 1.0, 12, 9, 1),
    (150, 23, 0.9, 14, 6, 1),
    (300, 4, 1.0, 10, 32, 1),
    (150, 0, 1.0, 6, 16, 1),
    (257, 0, 0.7, 2, 2, 2),
]


@pytest.mark.parametrize("L,K,jz,lines,slabs,span", CASES)
def test_stencil_col_hv_parity(L, K, jz, lines, slabs, span):
    print(f"[gpu] stencil column H.v L={L} K={K} lines={lines} slabs={slabs} span={span}", flush=True)
    q3 = _q3()
    bonds = orc.chain_bonds(L, [(1, 1.0, jz)])
    om = orc.Model(L, L, 1, 1, bonds)
    rk = om.rk_from_K(K)
    osec = om.build_sector(4, rk)
    gm = q3.Model(L, L, 1, 1, bonds)
    try:
        _opts(q3, 1, lines, slabs, span)
        s = q3.Sector.build(gm, 4, rk)
        rng = np.random.default_rng(3 * L + K)
        v = rng.uniform(-1, 1, s.dim) + 1j * rng.uniform(-1, 1, s.dim)
        ref = om.hv(osec, v)
        got = s.ham_to_vec_wave_vector(v)
        assert s.hv_last_launch()[0] == 5
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
        ref32 = ref - om.fm_energy() * v
        outs = {}
        for kern in (2, 1, 0):
            _opts(q3, kern, lines, slabs, span)
            for dt, prec in ((torch.complex128, q3.FP64), (torch.complex64, q3.FP32)):
                x = torch.from_numpy(v.astype(np.complex128)).to(dt).cuda()
                y = torch.full_like(x, float("nan"))
                s.hv_dev(y.data_ptr(), x.data_ptr(), prec)
                torch.cuda.synchronize()
                assert s.hv_last_launch()[0] == 5
                outs[(kern, prec)] = y.cpu().numpy()
        for prec in (q3.FP64, q3.FP32):
            assert np.array_equal(outs[(1, prec)], outs[(0, prec)]), prec
            assert np.array_equal(outs[(2, prec)], outs[(0, prec)]), prec
        got64 = outs[(1, q3.FP64)]
        assert np.max(np.abs(got64 - ref)) <= 1e-12 * np.max(np.abs(ref))
        got32 = outs[(1, q3.FP32)].astype(np.complex128)
        assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32))
    finally:
        _reset(q3)


@pytest.mark.parametrize("L,K,jz,lines,slabs,span", CASES[1:8])
def test_stencil_col_fused_lanczos(L, K, jz, lines, slabs, span):
    """The fused Lanczos step (option lanczos_fused 2: both precisions) with
    the column and pair kernels against the slab-window kernel.  Their H.v agree bit for
    bit, but alpha and beta come from per-block partial sums whose grouping
    follows each kernel's grid and thread-to-row map, so the runs agree to
    rounding: FP64 E0 to 1e-11 and the vectors to an overlap of 1 - 1e-9,
    FP32-mixed E0 to 1e-6 relative; and the FP64 E0 equals dense
    diagonalisation of the oracle's H to 1e-10 on small sectors."""
    print(f"[gpu] stencil column fused Lanczos L={L} K={K}", flush=True)
    q3 = _q3()
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, jz)]))
    sec = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
    kw = dict(maxitr=400, minitr=20, itrint=5, lnc_ene_conv=1e-13, seed=11)
    res = {}
    q3.set_option("lanczos_fused", 2)
    try:
        for kern in (2, 1, 0):
            _opts(q3, kern, lines, slabs, span)
            for prec in (q3.FP64, q3.FP32):
                E, psi, it = q3.lanczos_routines(sec, prec=prec, **kw)
                assert sec.hv_last_launch()[0] == 6
                res[(kern, prec)] = (E, psi, it)
    finally:
        q3.set_option("lanczos_fused", 1)
        _reset(q3)
    E1, p1, _ = res[(1, q3.FP64)]
    E0, p0, _ = res[(0, q3.FP64)]
    assert abs(E1 - E0) <= 1e-11 * max(1.0, abs(E0)), (E1, E0)
    assert abs(abs(np.vdot(p0, p1)) - 1.0) <= 1e-9
    E0, p0, _ = res[(0, q3.FP32)]
    for kern in (1, 2):
        E1, p1, _ = res[(kern, q3.FP32)]
        assert abs(E1 - E0) <= 1e-6 * max(1.0, abs(E0)), (kern, E1, E0)
    if sec.dim <= 3000:
        om = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, jz)]))
        osec = om.build_sector(4, om.rk_from_K(K))
        w = np.linalg.eigvalsh(om.full_hamiltonian(osec))
        assert abs(res[(1, q3.FP64)][0] - w[0]) <= 1e-10


def test_stencil_col_c3_bit_identical():
    """L = 1000, N_down = 4, k = 0 and k = 2 pi 7 / 1000 (41,417,250 rows):
    column, pair and slab-window H.v agree bit for bit in both precisions (the
    pair form runs in FP32), and the fused FP32 plain Lanczos with the pair
    kernel reaches the Bethe-ansatz E0 within 1e-9."""
    q3 = _q3()
    L = 1000
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 1.0)]))
    try:
        for K in (0, 7):
            print(f"[gpu] stencil column vs slab window L=1000 K={K}", flush=True)
            s = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
            g = torch.Generator(device="cuda")
            g.manual_seed(17 + K)
            for dt, prec in ((torch.complex128, q3.FP64), (torch.complex64, q3.FP32)):
                x = torch.randn(s.dim, dtype=dt, device="cuda", generator=g)
                y0 = torch.full_like(x, float("nan"))
                _opts(q3, 0)
                s.hv_dev(y0.data_ptr(), x.data_ptr(), prec)
                for kern in (1, 2):
                    y1 = torch.full_like(x, float("nan"))
                    _opts(q3, kern)
                    s.hv_dev(y1.data_ptr(), x.data_ptr(), prec)
                    torch.cuda.synchronize()
                    assert torch.equal(y1, y0), (K, prec, kern)
                    del y1
                del x, y0
            if K == 0:
                _opts(q3, 2)
                q3.set_option("lanczos_fused", 2)
                try:
                    E32, _, _ = q3.lanczos_routines(s, maxitr=1500, minitr=20, itrint=5, lnc_ene_conv=1e-11,
                                                    prec=q3.FP32, seed=1, want_vector=False)
                    assert s.hv_last_launch()[0] == 6
                finally:
                    q3.set_option("lanczos_fused", 1)
                assert abs(E32 / 242.000099091345514 - 1.0) <= 1e-9, E32
            del s
    finally:
        _reset(q3)
