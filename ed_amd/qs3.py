"""Host-side mirror of QS3's module interfaces over libqs3x (ctypes).

The names follow the reference's Fortran modules so that callers and tests
read like the reference:

  state_lists  -> Model, Sector.build (allocate_lists_omp_SQ), Sector.lists
  ham2vec      -> Sector.ham_to_vec_wave_vector
  lanczos      -> lanczos_routines
  eigen_solver -> my_trlanczos_routines_complex, calcu_FM_energy,
                  Sector.full_diag (Full_diag_routines, ALG = 3)
  ham2vec      -> Sector.make_full_hamiltonian (make_full_hamiltonian)
  get_expectation_values -> get_cf_2_wave_vector, get_lm_2_wave_vector,
                  calcu_Sq_a, calcu_DSF_wavevector
  DSFLan       -> dsf_analysis

Every call goes to the HIP library; there is no CPU fallback.  Importing this
module loads ed_amd/libqs3x.so and raises if it is missing.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqs3x.so")

FP64 = 0
FP32 = 1

_i32, _i64, _u64, _dbl, _P = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p


class QS3XError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise QS3XError(f"libqs3x.so not built ({LIB_PATH}); run ed_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    L.qs3x_last_error.restype = ctypes.c_char_p
    L.qs3x_wavevector.restype = _dbl
    L.qs3x_wavevector.argtypes = [_i32, _i32]
    sigs = {
        "qs3x_set_device": [_i32],
        "qs3x_set_option": [ctypes.c_char_p, _dbl],
        "qs3x_model_create": [_i32, _i32, _i32, _i32, _i32, _P, _P, _P, _P],
        "qs3x_model_destroy": [_P],
        "qs3x_fm_energy": [_P, _P],
        "qs3x_sector_build": [_P, _i32, _dbl, _dbl, _dbl, _P, _P],
        "qs3x_sector_from_lists": [_P, _i32, _i64, _P, _P, _P, _P],
        "qs3x_sector_build_u1": [_P, _i32, _P, _P],
        "qs3x_sector_destroy": [_P],
        "qs3x_sector_dim": [_P, _P],
        "qs3x_sector_lists": [_P, _P, _P],
        "qs3x_sector_nnz": [_P, _P],
        "qs3x_ham_to_vec_wave_vector": [_P, _P, _P],
        "qs3x_ham_to_vec_wave_vector_noalloc": [_P, _P, _P],
        "qs3x_hv_dev": [_P, _i32, _P, _P, _P],
        "qs3x_make_full_hamiltonian": [_P, _P, _i64],
        "qs3x_full_diag": [_P, _P, _P, _i64],
        "qs3x_time_hv": [_P, _i32, _i32, _P],
        "qs3x_hv_last_launch": [_P, _P, _P],
        "qs3x_hv_stencil_info": [_P, _i32, _P],
        "qs3x_hv_autotune": [_P, _i32, _i32, _P, _P, _i32, _P, _P],
        "qs3x_hv_tune_clear": [_P],
        "qs3x_block_tune_last": [_i32, _P, _P],
        "qs3x_lanczos": [_P, _i32, _i64, _i64, _i64, _dbl, _P, _u64, _P, _P, _P],
        "qs3x_trlanczos": [_P, _i32, _i32, _i32, _i32, _i64, _dbl, _P, _u64, _P, _P, _P],
        "qs3x_solver_trace": [_P, _i64, _P],
        "qs3x_solver_certificate": [_P],
        "qs3x_sector_build_all_k": [_P, _i32, _P],
        "qs3x_lanczos_group": [_P, _i32, _i32, _i64, _i64, _i64, _dbl, _u64, _i32, _P, _P, _P],
        "qs3x_dsf_sminus_batch": [_P, _i32, _P, _P, _P, _i32, _dbl, _i32, _P, _P, _P, _P],
        "qs3x_correlations": [_P, _P, _i64, _P, _P, _P],
        "qs3x_local_mag": [_P, _P, _P],
        "qs3x_apply_sq": [_P, _P, _i32, _P, _P, _P, _P],
        "qs3x_dsf_recurrence": [_P, _i32, _i32, _P, _i32, _dbl, _P, _P, _P],
        "qs3x_continued_fraction": [_i32, _P, _P, _dbl, _dbl, _dbl, _i64, _P, _i32, _P],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    return L


_lib = _load()


def lib():
    return _lib


def _check(rc):
    if rc != 0:
        raise QS3XError(_lib.qs3x_last_error().decode())


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_option(name, value):
    """qs3x_set_option: 'tr_reorth' (0 full, 1 locked + local), 'tr_cgs_passes', 'fp32_floor',
    'hv_kernel', 'verbose'."""
    _check(_lib.qs3x_set_option(name.encode(), float(value)))


def wavevector(K, L):
    """rk = 2 acos(-1) K / L, as read_ip evaluates it (input_param.f90:76-78)."""
    return _lib.qs3x_wavevector(int(K), int(L))


class Model:
    """Lattice and XXZ bond list (read_ip, input_param.f90:58-103).

    bonds: rows (i, j, Jxy, Jz) in bond-file order, 1-based sites."""

    def __init__(self, NOS, LX, LY, LZ, bonds):
        b = np.asarray(bonds, dtype=np.float64).reshape(-1, 4)
        self.NOS, self.LX, self.LY, self.LZ = int(NOS), int(LX), int(LY), int(LZ)
        self.bonds = b
        pairs = np.ascontiguousarray(b[:, :2].astype(np.int32).reshape(-1))
        jxy = np.ascontiguousarray(b[:, 2])
        jz = np.ascontiguousarray(b[:, 3])
        h = ctypes.c_void_p()
        _check(_lib.qs3x_model_create(self.NOS, self.LX, self.LY, self.LZ, len(b), _p(pairs),
                                      _p(jxy), _p(jz), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if self._h:
                _lib.qs3x_model_destroy(self._h)
        except Exception:
            pass

    def calcu_FM_energy(self):
        """eigen_solver.f90:341-347."""
        e = np.zeros(1)
        _check(_lib.qs3x_fm_energy(self._h, _p(e)))
        return float(e[0])

    def rk_from_K(self, KX, KY=0, KZ=0):
        return (wavevector(KX, self.LX), wavevector(KY, self.LY), wavevector(KZ, self.LZ))


class Sector:
    """An (N_down, k) sector on the device (allocate_lists_omp_SQ)."""

    def __init__(self, model, nod, handle, rk=None):
        self.model = model
        self.nod = int(nod)
        self.rk = rk
        self._h = handle
        d = np.zeros(1, dtype=np.int64)
        _check(_lib.qs3x_sector_dim(self._h, _p(d)))
        self.dim = int(d[0])

    @classmethod
    def build(cls, model, nod, rk):
        h = ctypes.c_void_p()
        d = np.zeros(1, dtype=np.int64)
        _check(_lib.qs3x_sector_build(model._h, int(nod), float(rk[0]), float(rk[1]), float(rk[2]),
                                      ctypes.byref(h), _p(d)))
        return cls(model, nod, h, rk)

    @classmethod
    def build_u1(cls, model, nod):
        """qs3x_sector_build_u1: the U(1)-only block of every configuration
        with nod down spins (source_only_u1), rows in colex rank order."""
        h = ctypes.c_void_p()
        d = np.zeros(1, dtype=np.int64)
        _check(_lib.qs3x_sector_build_u1(model._h, int(nod), ctypes.byref(h), _p(d)))
        return cls(model, nod, h)

    @classmethod
    def build_all_k(cls, model, nod):
        """Every translation representative of N_down (the union of all k
        sectors of a chain), for the batched multi-k entry points."""
        h = ctypes.c_void_p()
        _check(_lib.qs3x_sector_build_all_k(model._h, int(nod), ctypes.byref(h)))
        return cls(model, nod, h)

    @classmethod
    def from_lists(cls, model, nod, list_s, list_r, explist=None):
        ls = np.ascontiguousarray(list_s, dtype=np.int64)
        lr = np.ascontiguousarray(list_r, dtype=np.float64)
        ex = None if explist is None else np.ascontiguousarray(explist, dtype=np.complex128)
        h = ctypes.c_void_p()
        _check(_lib.qs3x_sector_from_lists(model._h, int(nod), len(ls), _p(ls), _p(lr), _p(ex),
                                           ctypes.byref(h)))
        return cls(model, nod, h)

    def __del__(self):
        try:
            if self._h:
                _lib.qs3x_sector_destroy(self._h)
        except Exception:
            pass

    def lists(self):
        ls = np.zeros(self.dim, dtype=np.int64)
        lr = np.zeros(self.dim, dtype=np.float64)
        _check(_lib.qs3x_sector_lists(self._h, _p(ls), _p(lr)))
        return ls, lr

    def nnz(self):
        n = np.zeros(1, dtype=np.int64)
        _check(_lib.qs3x_sector_nnz(self._h, _p(n)))
        return int(n[0])

    def ham_to_vec_wave_vector(self, v1):
        """v0 = H v1 (ham2vec_v3.f90:37-72), host vectors."""
        a = _c128(v1)
        assert a.shape == (self.dim,)
        out = np.zeros_like(a)
        _check(_lib.qs3x_ham_to_vec_wave_vector(self._h, _p(out), _p(a)))
        return out

    def ham_to_vec_wave_vector_noalloc(self, v1):
        """v0 = H v1 (host vectors) staged through the library's static device
        array: no device allocation; dim <= 2,097,152."""
        a = _c128(v1)
        assert a.shape == (self.dim,)
        out = np.zeros_like(a)
        _check(_lib.qs3x_ham_to_vec_wave_vector_noalloc(self._h, _p(out), _p(a)))
        return out

    def make_full_hamiltonian(self):
        """The dense H (ham2vec_v3.f90:5-35): H[i, j] = row i's term on
        configuration j, built as one H e_j per column on the device."""
        h = np.zeros((self.dim, self.dim), dtype=np.complex128, order="F")
        _check(_lib.qs3x_make_full_hamiltonian(self._h, _p(h), self.dim))
        return h

    def full_diag(self, vectors=True):
        """Full diagonalisation (Full_diag_routines, eigen_solver.f90:51-67,
        263-339): eigenvalues ascending and, with vectors, the orthonormal
        eigenvectors as the columns of a (dim, dim) array."""
        w = np.zeros(self.dim)
        v = np.zeros((self.dim, self.dim), dtype=np.complex128, order="F") if vectors else None
        _check(_lib.qs3x_full_diag(self._h, _p(w), _p(v), self.dim))
        return (w, v) if vectors else w

    def hv_dev(self, v0_ptr, v1_ptr, prec=FP64, stream=None):
        """Device-pointer H.v (FP32: v0 = (H - e_zz) v1)."""
        _check(_lib.qs3x_hv_dev(self._h, int(prec), ctypes.c_void_p(v0_ptr), ctypes.c_void_p(v1_ptr),
                                ctypes.c_void_p(stream or 0)))

    def hv_last_launch(self):
        """(kernel family, dynamic LDS bytes) of the last H.v launch: 0
        nearest-neighbour chain, 1 batched-load chain, 2 per-hop chain, 3
        general lattice, 4 U(1)-only, 5 gap-lattice stencil, 6 the fused
        Lanczos step on the stencil."""
        f = np.zeros(1, dtype=np.int32)
        lds = np.zeros(1, dtype=np.int64)
        _check(_lib.qs3x_hv_last_launch(self._h, _p(f), _p(lds)))
        return int(f[0]), int(lds[0])

    def hv_stencil_info(self, prec=FP32):
        """Plan of the stencil H.v (after its first launch): dict of slow_rows,
        lines, tasks, tile_lines, slabs, grid, form (0 slab window, 1 column,
        2 pair) and span (window positions per lane of the column form)."""
        info = np.zeros(8, dtype=np.int64)
        _check(_lib.qs3x_hv_stencil_info(self._h, int(prec), _p(info)))
        return dict(zip(["slow_rows", "lines", "tasks", "tile_lines", "slabs", "grid", "form", "span"],
                        info.tolist()))

    def time_hv(self, prec=FP64, n_rep=10):
        ms = np.zeros(1)
        _check(_lib.qs3x_time_hv(self._h, int(prec), int(n_rep), _p(ms)))
        return float(ms[0])

    def hv_autotune(self, prec=FP64, n_rep=5):
        """qs3x_hv_autotune: times the kernel families, then the winner's own
        parameters, and keeps the fastest for this sector and precision.
        Returns (best dict, best ms, array of rows (family, kernel, stencil,
        lines, slabs, depth, pairs, chunk, xcd, quarters, form, ms))."""
        best = np.zeros(11, dtype=np.int32)
        ms = np.zeros(1)
        res = np.zeros((128, 12))
        n = np.zeros(1, dtype=np.int32)
        _check(_lib.qs3x_hv_autotune(self._h, int(prec), int(n_rep), _p(best), _p(ms), 128, _p(n), _p(res)))
        keys = ("family", "kernel", "stencil", "lines", "slabs", "depth", "pairs", "chunk", "xcd", "grid_quarters",
                "form")
        return dict(zip(keys, (int(x) for x in best))), float(ms[0]), res[:int(n[0])].copy()

    def hv_tune_clear(self):
        _check(_lib.qs3x_hv_tune_clear(self._h))


# --------------------------------------------------------------- solvers
def lanczos_routines(sector, maxitr=2000, minitr=20, itrint=5, lnc_ene_conv=1e-14, prec=FP64,
                     start=None, seed=1, want_vector=True):
    """lanczos_routines / lancz_eigen_val (lanczos.f90:14-152) on the device.
    Returns (E0, psi or None, iterations)."""
    st = None if start is None else _c128(start)
    ene = np.zeros(2)
    psi = np.zeros(sector.dim, dtype=np.complex128) if want_vector else None
    it = np.zeros(1, dtype=np.int64)
    _check(_lib.qs3x_lanczos(sector._h, int(prec), int(maxitr), int(minitr), int(itrint),
                             float(lnc_ene_conv), _p(st), int(seed), _p(ene), _p(psi), _p(it)))
    return float(ene[0]), psi, int(it[0])


def lanczos_group(sectors, maxitr=2000, minitr=20, itrint=5, lnc_ene_conv=1e-14, prec=FP64, seed=1,
                  use_graph=True):
    """Plain Lanczos energies of many sectors of one chain model at once
    (lanczos_routines per sector, lanczos.f90:14-152).  Returns
    (E0 per sector, iterations per sector, stats dict)."""
    hs = (ctypes.c_void_p * len(sectors))(*[s._h.value if isinstance(s._h, ctypes.c_void_p) else s._h
                                           for s in sectors])
    ene = np.zeros(len(sectors))
    its = np.zeros(len(sectors), dtype=np.int64)
    stats = np.zeros(8)
    _check(_lib.qs3x_lanczos_group(hs, len(sectors), int(prec), int(maxitr), int(minitr), int(itrint),
                                   float(lnc_ene_conv), int(seed), 1 if use_graph else 0, _p(ene), _p(its),
                                   _p(stats)))
    keys = ["seconds", "seconds_setup", "steps", "step_launches", "rows_padded", "groups", "hop_cache_bytes"]
    return ene, its, dict(zip(keys, stats[:7].tolist()))


def my_trlanczos_routines_complex(sector, NOE, NOK, NOM, maxitr=1000, lnc_ene_conv=1e-14, prec=FP64,
                                  start=None, seed=1, want_vectors=True):
    """my_trlanczos_routines_complex (eigen_solver.f90:10-261) on the device.
    Returns (ene[NOE], psi[dim, NOE] or None, stats dict)."""
    st = None if start is None else _c128(start)
    ene = np.zeros(NOE)
    psi = np.zeros((NOE, sector.dim), dtype=np.complex128) if want_vectors else None
    stats = np.zeros(8)
    _check(_lib.qs3x_trlanczos(sector._h, int(prec), int(NOE), int(NOK), int(NOM), int(maxitr),
                               float(lnc_ene_conv), _p(st), int(seed), _p(ene), _p(psi), _p(stats)))
    keys = ["restarts", "hv_fp32", "hv_fp64", "seconds", "seconds_fp32", "max_residual", "rr_accepted",
            "seconds_alloc"]
    return ene, (psi.T.copy() if psi is not None else None), dict(zip(keys, stats.tolist()))


def solver_certificate():
    """Certificate of the last FP32-mixed thick restart: dict of gap (spectral
    gap estimate), residual (FP64), bound (r^2 / gap), accepted (1 certified,
    2 agreement only, 0 not accepted, -1 none), and the last thick restart's
    seconds allocating and freeing its Krylov bases."""
    out = np.zeros(6)
    _check(_lib.qs3x_solver_certificate(_p(out)))
    return dict(zip(["gap", "residual", "bound", "accepted", "seconds_alloc", "seconds_free"], out.tolist()))


def solver_trace():
    """Convergence trace of the last solve: array of rows
    (seconds since start, H.v count, lowest Ritz value, precision 32/64)."""
    n = np.zeros(1, dtype=np.int64)
    _check(_lib.qs3x_solver_trace(ctypes.c_void_p(0), 0, _p(n)))
    out = np.zeros((int(n[0]), 4))
    if n[0]:
        _check(_lib.qs3x_solver_trace(_p(out), int(n[0]), _p(n)))
    return out


def get_cf_2_wave_vector(sector, psi, pairs):
    """<SzSz>, <S+S-> for the FILECF pair list (get_expectation_values.f90:234-279).
    pairs: rows (a, b) as the file lines read."""
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1))
    v = _c128(psi)
    szsz = np.zeros(len(pr) // 2)
    spsm = np.zeros(len(pr) // 2, dtype=np.complex128)
    _check(_lib.qs3x_correlations(sector._h, _p(v), len(pr) // 2, _p(pr), _p(szsz), _p(spsm)))
    return szsz, spsm


def get_lm_2_wave_vector(sector, psi):
    """<S^z_j> for every site (get_expectation_values.f90:197-232)."""
    v = _c128(psi)
    out = np.zeros(sector.model.NOS)
    _check(_lib.qs3x_local_mag(sector._h, _p(v), _p(out)))
    return out


def calcu_Sq_a(spsmsz, gs, tgt, q, pos, psi):
    """S^a_q |psi> (get_expectation_values.f90:281-377), pull form on the device."""
    qv = np.ascontiguousarray(q, dtype=np.float64)
    ps = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    v = _c128(psi)
    f = np.zeros(tgt.dim, dtype=np.complex128)
    _check(_lib.qs3x_apply_sq(gs._h, tgt._h, int(spsmsz), _p(qv), _p(ps), _p(v), _p(f)))
    return f


def dsf_recurrence(tgt, f, itr_dsf, rfield=0.0, prec=FP64):
    """The Lanczos recurrence of calcu_DSF_wavevector (get_expectation_values.f90:481-592).
    f: (dim,) or (nvec, dim).  Returns alpha, beta2 (beta2[...,0] = 0), norm1."""
    F = np.atleast_2d(_c128(f))
    nvec = F.shape[0]
    alpha = np.zeros((nvec, itr_dsf))
    beta2 = np.zeros((nvec, itr_dsf))
    norm1 = np.zeros(nvec)
    Fc = np.ascontiguousarray(F)
    _check(_lib.qs3x_dsf_recurrence(tgt._h, int(prec), nvec, _p(Fc), int(itr_dsf), float(rfield),
                                    _p(alpha), _p(beta2), _p(norm1)))
    if np.ndim(f) == 1:
        return alpha[0], beta2[0], float(norm1[0])
    return alpha, beta2, norm1


def dsf_sminus_batch(gs, K, pos, psi, itr_dsf, rfield=0.0, prec=FP32):
    """S^-_q |psi> and the DSF recurrence for up to 64 wave numbers K (target
    sectors k = 2 pi K / L of N_down + 1) in one block.  Returns alpha, beta2
    (len(K) x itr_dsf), norm1 (len(K)) and a stats dict."""
    Ks = np.ascontiguousarray(np.asarray(K, dtype=np.int32).reshape(-1))
    nq = len(Ks)
    ps = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    v = _c128(psi)
    alpha = np.zeros((nq, itr_dsf))
    beta2 = np.zeros((nq, itr_dsf))
    norm1 = np.zeros(nq)
    stats = np.zeros(8)
    _check(_lib.qs3x_dsf_sminus_batch(gs._h, nq, _p(Ks), _p(ps), _p(v), int(itr_dsf), float(rfield), int(prec),
                                      _p(alpha), _p(beta2), _p(norm1), _p(stats)))
    keys = ["seconds_setup", "seconds_sminus", "seconds_recurrence", "ms_per_block_step", "dim_all_k", "steps",
            "block_stencil", "rows_per_row_kernel"]
    return alpha, beta2, norm1, dict(zip(keys, stats[:8].tolist()))


def block_tune_last():
    """qs3x_block_tune_last: rows (lines, threads, slabs, ms) of the last
    batched DSF block-stencil tuning (option block_autotune), the chosen
    shape last; empty without it."""
    rows = np.zeros((32, 4))
    n = np.zeros(1, dtype=np.int32)
    _check(_lib.qs3x_block_tune_last(32, _p(n), _p(rows)))
    return rows[:int(n[0])].copy()


def dsf_analysis(alpha, beta2, norm1, E0, omega, eta=0.04, reference_quirk=True):
    """S(q, w) continued fraction (source_DSFLan/main.f90:40-94)."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta2, dtype=np.float64)
    w = np.ascontiguousarray(omega, dtype=np.float64)
    out = np.zeros(len(w))
    _check(_lib.qs3x_continued_fraction(len(a), _p(a), _p(b), float(norm1), float(E0), float(eta),
                                        len(w), _p(w), 1 if reference_quirk else 0, _p(out)))
    return out


def chain_bonds(L, couplings):
    """Periodic chain bond list: couplings [(d, Jxy, Jz), ...] -> bonds (i, i+d)."""
    out = []
    for d, jxy, jz in couplings:
        for i in range(1, L + 1):
            out.append((i, (i - 1 + d) % L + 1, jxy, jz))
    return out


def start_vector(dim, seed, row0=0):
    """The engine's counter-based start vector (splitmix64 of seed and row):
    re, im uniform in [-1, 1).  Same values as the device generator."""
    M = (1 << 64) - 1
    r = np.arange(row0, row0 + dim, dtype=np.uint64)

    def mix(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(M)
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(M)
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(M)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        k = mix(np.uint64(seed) ^ (r * np.uint64(0xD1B54A32D192ED03)))
        a = mix(k)
        b = mix(a)
    re = (a >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0
    im = (b >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0
    return re + 1j * im
