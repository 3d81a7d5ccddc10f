"""ctypes bindings and numpy solver restatements for the QS3 oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Parity status of each oracle function (how it is pinned, tests/ file):
  binomials, rank/unrank, j_flip, representative, checkstate, sector build,
  H.v, calcu_cf_trans_2, calcu_lm_trans_2, calcu_Sq_a
      -- pinned bit-exactly (integers, list_r) or to 1e-14 (H.v, observables)
         against the reference's own Fortran modules (oracle/_ref) on chains
         and 2D tori, k = 0 and k != 0 (tests/test_oracle_vs_reference.py);
         also by brute-force orbit enumeration, the character formula for
         sector dimensions and dense symmetry-free Hamiltonians
         (tests/test_oracle_math.py).
  build_sector_chain
      -- pinned against build_sector (literal all-rank scan) on small chains.
  lanczos_plain, trlanczos
      -- the reference's lanczos/eigen_solver modules need LAPACK and cannot
         be built here; pinned by exact spectra (ferromagnet, one magnon,
         XX free fermions, Bethe ansatz) and dense numpy.linalg.eigh
         (tests/test_oracle_math.py).
  dsf_recurrence
      -- pinned by the sum rule norm1 = <psi|S+_q S-_q|psi> (through the
         correlation functions), by the spectral (Lehmann) representation of
         the resolvent on small dense sectors, and by the first moment alpha_1
         (tests/test_oracle_math.py).
  continued_fraction
      -- reference_quirk=True: pinned to 1e-13 against the output of the
         reference's own DSFLan program compiled here (oracle/_ref/dsflan.exe;
         fixtures tests/golden/dsflan_*.npz, tests/test_dsflan_reference.py);
         reference_quirk=False: pinned by the Lehmann sum of the Jacobi matrix.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_qs3.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libqs3ref.so")
_REF_U1_PATH = os.path.join(_HERE, "_ref", "libqs3ref_u1.so")

i64 = ctypes.c_int64
dbl = ctypes.c_double
P = ctypes.c_void_p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run oracle/build_oracle.sh")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_model_new.restype = P
        L.orc_model_new.argtypes = [i64, i64, i64, i64, i64, P, P, P, P, i64]
        L.orc_model_free.argtypes = [P]
        L.orc_f_combination.restype = i64
        L.orc_f_combination.argtypes = [i64, i64]
        L.orc_inv_list.restype = i64
        L.orc_inv_list.argtypes = [P, P, i64]
        L.orc_list_fly.argtypes = [P, i64, i64, P]
        L.orc_j_flip.argtypes = [i64, i64, P, i64, P]
        L.orc_findstate.restype = i64
        L.orc_findstate.argtypes = [i64, P, i64]
        L.orc_representative.restype = i64
        L.orc_representative.argtypes = [P, i64, P, P]
        L.orc_checkstate.restype = dbl
        L.orc_checkstate.argtypes = [P, i64, i64, dbl, dbl, dbl, P]
        L.orc_build_sector.restype = i64
        L.orc_build_sector.argtypes = [P, i64, dbl, dbl, dbl, P, P, P, P]
        L.orc_build_sector_chain.restype = i64
        L.orc_build_sector_chain.argtypes = [P, i64, dbl, dbl, dbl, P, P]
        L.orc_ham_to_vec.argtypes = [P, i64, i64, P, P, P, P, P, i64, i64]
        L.orc_u1_ham_to_vec.argtypes = [P, i64, P, P, i64, i64]
        L.orc_count_hops.restype = i64
        L.orc_count_hops.argtypes = [P, i64, i64, P, i64, i64]
        L.orc_mk_list_ij.argtypes = [P, P, P, P]
        L.orc_calcu_cf_trans_2.argtypes = [P, i64, i64, P, P, P, P, P, i64, P, P]
        L.orc_calcu_lm_trans_2.restype = dbl
        L.orc_calcu_lm_trans_2.argtypes = [P, i64, i64, P, P, P, i64]
        L.orc_mk_list_n.argtypes = [P, i64, P, P]
        L.orc_calcu_Sq_a.argtypes = [P, i64, i64, i64, P, P, i64, i64, P, P, P, P, P, P]
        L.orc_full_hamiltonian.argtypes = [P, i64, i64, P, P, P, P]
        L.orc_shift_tables.argtypes = [P, P, P, P]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def wavevector(K, L):
    """rk = 2 pi K / L exactly as read_ip evaluates it (input_param.f90:76-78)."""
    PI = math.acos(-1.0)
    return 2.0 * PI * float(K) / float(L)


def explist(LX, LY, LZ, rkx, rky, rkz):
    """explist(0:LX,0:LY,0:LZ) = exp(i(rkx i + rky j + rkz k)) (input_param.f90:88-89),
    flattened column-major.  Evaluated as cos/sin of the same argument, which is
    what cexp(0 + i x) returns."""
    out = np.empty((LZ + 1, LY + 1, LX + 1), dtype=np.complex128)
    for k in range(LZ + 1):
        for j in range(LY + 1):
            for i in range(LX + 1):
                x = rkx * float(i) + rky * float(j) + rkz * float(k)
                out[k, j, i] = complex(math.cos(x), math.sin(x))
    return out.reshape(-1)


class Sector:
    def __init__(self, nod, rk, list_s, list_r, ex):
        self.nod = nod
        self.rk = rk
        self.list_s = list_s
        self.list_r = list_r
        self.explist = ex

    @property
    def dim(self):
        return len(self.list_s)


class Model:
    """Lattice + bond list.  bonds: sequence of (i, j, Jxy, Jz), 1-based sites,
    in bond-file order (list_xxz_term_*.dat)."""

    def __init__(self, NOS, LX, LY, LZ, bonds, ncol=None):
        b = np.asarray(bonds, dtype=np.float64).reshape(-1, 4)
        self.NOS, self.LX, self.LY, self.LZ = int(NOS), int(LX), int(LY), int(LZ)
        self.bonds = b
        self.pi = np.ascontiguousarray(b[:, 0].astype(np.int64))
        self.pj = np.ascontiguousarray(b[:, 1].astype(np.int64))
        self.jxy = np.ascontiguousarray(b[:, 2])
        self.jz = np.ascontiguousarray(b[:, 3])
        self.ncol = int(ncol) if ncol else min(self.NOS, 8)
        self._h = lib().orc_model_new(self.NOS, self.LX, self.LY, self.LZ, len(b),
                                      _ptr(self.pi), _ptr(self.pj), _ptr(self.jxy),
                                      _ptr(self.jz), self.ncol)

    def __del__(self):
        try:
            if self._h:
                lib().orc_model_free(self._h)
        except Exception:
            pass

    # -- combinatorics -------------------------------------------------------
    def comb(self, n, k):
        return lib().orc_f_combination(n, k)

    def rank(self, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.int64))
        return lib().orc_inv_list(self._h, _ptr(a), len(a))

    def unrank(self, t, nod):
        out = np.zeros(nod, dtype=np.int64)
        lib().orc_list_fly(self._h, int(t), nod, _ptr(out))
        return out

    def j_flip(self, i, j, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.int64))
        out = np.zeros(len(a), dtype=np.int64)
        lib().orc_j_flip(i, j, _ptr(a), len(a), _ptr(out))
        return out

    def representative(self, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.int64)).copy()
        ell = np.zeros(3, dtype=np.int64)
        t = lib().orc_representative(self._h, len(a), _ptr(a), _ptr(ell))
        return t, ell, a

    def checkstate(self, s, nod, rk):
        c = np.zeros(1, dtype=np.int64)
        ro = lib().orc_checkstate(self._h, int(s), nod, rk[0], rk[1], rk[2], _ptr(c))
        return ro, int(c[0])

    def shift_tables(self):
        n = self.NOS
        sx, sy, sz = (np.zeros(n, dtype=np.int64) for _ in range(3))
        lib().orc_shift_tables(self._h, _ptr(sx), _ptr(sy), _ptr(sz))
        return sx, sy, sz

    # -- sectors -------------------------------------------------------------
    def rk_from_K(self, KX, KY=0, KZ=0):
        return (wavevector(KX, self.LX), wavevector(KY, self.LY), wavevector(KZ, self.LZ))

    def build_sector(self, nod, rk):
        """allocate_lists_omp_SQ (state_lists.f90:225-292), literal scan."""
        L = lib()
        THS = self.comb(self.NOS, nod) if nod > 0 else 1
        flags = np.zeros(THS, dtype=np.uint8)
        ro = np.zeros(THS, dtype=np.float64)
        D = L.orc_build_sector(self._h, nod, rk[0], rk[1], rk[2], None, None, _ptr(flags), _ptr(ro))
        ls = np.zeros(D, dtype=np.int64)
        lr = np.zeros(D, dtype=np.float64)
        L.orc_build_sector(self._h, nod, rk[0], rk[1], rk[2], _ptr(ls), _ptr(lr), _ptr(flags), _ptr(ro))
        return Sector(nod, rk, ls, lr, explist(self.LX, self.LY, self.LZ, *rk))

    def build_sector_chain(self, nod, rk):
        L = lib()
        D = L.orc_build_sector_chain(self._h, nod, rk[0], rk[1], rk[2], None, None)
        if D < 0:
            raise ValueError("build_sector_chain needs LY = LZ = 1")
        ls = np.zeros(D, dtype=np.int64)
        lr = np.zeros(D, dtype=np.float64)
        L.orc_build_sector_chain(self._h, nod, rk[0], rk[1], rk[2], _ptr(ls), _ptr(lr))
        return Sector(nod, rk, ls, lr, explist(self.LX, self.LY, self.LZ, *rk))

    # -- operators -----------------------------------------------------------
    def hv(self, sec, v, rows=None):
        """ham_to_vec_wave_vector (ham2vec_v3.f90:37-72)."""
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1)
        r0, r1 = (0, sec.dim) if rows is None else rows
        lib().orc_ham_to_vec(self._h, sec.nod, sec.dim, _ptr(sec.list_s), _ptr(sec.list_r),
                             _ptr(sec.explist), _ptr(v0), _ptr(v1), r0, r1)
        return v0

    def u1_dim(self, nod):
        """THS of the U(1)-only code: combination(NOS, NOD) (source_only_u1/main.f90:27)."""
        return int(lib().orc_f_combination(self.NOS, nod))

    def hv_u1(self, nod, v, rows=None):
        """U(1)-only ham_to_vec_wave_vector (source_only_u1/ham2vec_v3.f90:35-65):
        rows are colex ranks 1..C(NOS, nod).  The model needs ncol >= nod."""
        if self.ncol < nod:
            raise ValueError("hv_u1 needs a model with ncol >= nod")
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1)
        r0, r1 = (0, len(v1)) if rows is None else rows
        lib().orc_u1_ham_to_vec(self._h, nod, _ptr(v0), _ptr(v1), r0, r1)
        return v0

    def count_hops(self, sec, rows=None):
        r0, r1 = (0, sec.dim) if rows is None else rows
        return lib().orc_count_hops(self._h, sec.nod, sec.dim, _ptr(sec.list_s), r0, r1)

    def full_hamiltonian(self, sec):
        H = np.zeros((sec.dim, sec.dim), dtype=np.complex128)
        lib().orc_full_hamiltonian(self._h, sec.nod, sec.dim, _ptr(sec.list_s), _ptr(sec.list_r),
                                   _ptr(sec.explist), _ptr(H))
        return H

    def fm_energy(self):
        """calcu_FM_energy (eigen_solver.f90:341-347)."""
        return float(np.sum(self.jz)) / 4.0

    # -- observables ---------------------------------------------------------
    def mk_list_ij(self, r):
        rr = np.asarray(r, dtype=np.int64)
        out = np.zeros(2 * self.NOS, dtype=np.int64)
        n = np.zeros(1, dtype=np.int64)
        lib().orc_mk_list_ij(self._h, _ptr(rr), _ptr(out), _ptr(n))
        return out[: 2 * int(n[0])].reshape(-1, 2)

    def cf_pair(self, sec, v, r):
        """calcu_cf_trans_2 for the orbit of pair r = (i, j) as stored in list_ij."""
        lij = np.ascontiguousarray(self.mk_list_ij(r).reshape(-1))
        if len(lij) == 0:
            return 0, 0.0, 0j
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        szsz = np.zeros(1)
        spsm = np.zeros(1, dtype=np.complex128)
        lib().orc_calcu_cf_trans_2(self._h, sec.nod, sec.dim, _ptr(sec.list_s), _ptr(sec.list_r),
                                   _ptr(sec.explist), _ptr(v1), _ptr(lij), len(lij) // 2,
                                   _ptr(szsz), _ptr(spsm))
        return len(lij) // 2, float(szsz[0]), complex(spsm[0])

    def correlations(self, sec, v, pairs):
        """get_cf_2_wave_vector (get_expectation_values.f90:234-279): pairs is
        the FILECF list as read, i.e. list_expe_ss(2:1:-1,i) = (file col 2, col 1)
        stored as list_expe_ss(1,i) = col 2, list_expe_ss(2,i) = col 1.
        Returns (szsz[NOCF], spsm[NOCF]) indexed like the file lines."""
        NOS = self.NOS
        ez = np.zeros(NOS * NOS)
        es = np.zeros(NOS * NOS, dtype=np.complex128)
        lst = [(int(p[1]), int(p[0])) for p in pairs]   # (list_expe_ss(1,:), list_expe_ss(2,:))
        jmax = NOS * NOS
        no_tot = 0
        for j in range(jmax):
            if no_tot == jmax:
                break
            r = lst[j]
            lij = self.mk_list_ij(r)
            if len(lij) == 0:
                continue
            no_tot += len(lij)
            _, sz, sp = self.cf_pair(sec, v, r)
            for a, b in lij:
                ez[a - 1 + (b - 1) * NOS] = sz
                es[a - 1 + (b - 1) * NOS] = sp
        szsz = np.array([ez[r[0] - 1 + (r[1] - 1) * NOS] for r in lst])
        spsm = np.array([es[r[0] - 1 + (r[1] - 1) * NOS] for r in lst])
        return szsz, spsm

    def mk_list_n(self, n):
        out = np.zeros(self.NOS, dtype=np.int64)
        c = np.zeros(1, dtype=np.int64)
        lib().orc_mk_list_n(self._h, int(n), _ptr(out), _ptr(c))
        return out[: int(c[0])]

    def local_mag(self, sec, v):
        """get_lm_2_wave_vector (get_expectation_values.f90:197-232) for one vector."""
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        ez = np.zeros(self.NOS)
        no_tot = 0
        for j in range(1, self.NOS + 1):
            if no_tot == self.NOS:
                break
            ln = np.ascontiguousarray(self.mk_list_n(j))
            if len(ln) == 0:
                continue
            no_tot += len(ln)
            val = lib().orc_calcu_lm_trans_2(self._h, sec.nod, sec.dim, _ptr(sec.list_s),
                                             _ptr(v1), _ptr(ln), len(ln))
            ez[ln - 1] = val
        return ez

    def sq_a(self, spsmsz, sec, new, q, pos, psi):
        """calcu_Sq_a (get_expectation_values.f90:281-377), push form."""
        expks = np.array([np.exp(-1j * float(np.dot(q, pos[i]))) for i in range(self.NOS)],
                         dtype=np.complex128)
        fn = np.zeros(new.dim, dtype=np.complex128)
        v = np.ascontiguousarray(psi, dtype=np.complex128)
        lib().orc_calcu_Sq_a(self._h, spsmsz, sec.nod, sec.dim, _ptr(sec.list_s), _ptr(sec.list_r),
                             new.nod, new.dim, _ptr(new.list_s), _ptr(new.list_r), _ptr(new.explist),
                             _ptr(expks), _ptr(v), _ptr(fn))
        return fn


# ---------------------------------------------------------------------------
# Bond-list helpers (list_xxz_term_*.dat contents)
# ---------------------------------------------------------------------------
def chain_bonds(L, couplings):
    """Periodic chain; couplings = [(d, Jxy, Jz), ...] gives bonds (i, i+d mod L)
    for every i in site order, one family after the other."""
    out = []
    for d, jxy, jz in couplings:
        for i in range(1, L + 1):
            j = (i - 1 + d) % L + 1
            out.append((i, j, jxy, jz))
    return out


def square_bonds(LX, LY, J):
    out = []
    for y in range(LY):
        for x in range(LX):
            s = x + LX * y + 1
            out.append((s, (x + 1) % LX + LX * y + 1, J, J))
    for y in range(LY):
        for x in range(LX):
            s = x + LX * y + 1
            out.append((s, x + LX * ((y + 1) % LY) + 1, J, J))
    return out


def read_bond_file(path, n):
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                a = line.replace(",", " ").split()
                fl = [float(x.replace("d", "e").replace("D", "e")) for x in a[2:4]]
                rows.append((int(a[0]), int(a[1]), fl[0], fl[1]))
            if len(rows) == n:
                break
    return rows


# ---------------------------------------------------------------------------
# Solver restatements (numpy; small sectors)
# ---------------------------------------------------------------------------
def _tridiag_lowest(alpha, beta):
    """dstevx range 'I', il = iu = 1: lowest eigenpair of the real symmetric
    tridiagonal matrix (diag alpha, off-diagonal beta[:-1])."""
    n = len(alpha)
    T = np.diag(alpha) + np.diag(beta[: n - 1], 1) + np.diag(beta[: n - 1], -1)
    w, z = np.linalg.eigh(T)
    return w[0], z[:, 0]


def lanczos_plain(hv, v0, maxitr=1000, minitr=20, itrint=5, conv_tol=1e-14, want_vector=True):
    """lancz_eigen_val (lanczos.f90:50-152): three-term Lanczos without
    reorthogonalisation, lowest Ritz value every itrint steps past minitr,
    second pass for the eigenvector.  Returns (E0, psi, iterations, alpha, beta)."""
    dim = len(v0)
    psi1 = np.array(v0, dtype=np.complex128)
    b0 = np.linalg.norm(psi1)
    psi1 /= b0
    psi_io = psi1.copy()
    psi3 = np.zeros_like(psi1)
    alpha, beta = [], []
    tmp = 0.0
    mitr = min(maxitr, dim)
    ene, z, ite = None, None, 0
    for i in range(1, mitr + 1):
        psi2 = hv(psi1)
        a = float(np.vdot(psi1, psi2).real)
        alpha.append(a)
        if i > 1:
            psi2 = psi2 - a * psi1 - beta[-1] * psi3
        else:
            psi2 = psi2 - a * psi1
        b = float(np.linalg.norm(psi2))
        beta.append(b)
        psi3 = psi1
        psi1 = psi2 / b if b > 0 else psi2
        ite = i
        if i == mitr or (i >= minitr and i % itrint == 0):
            if i == 1:
                return alpha[0], np.eye(dim, 1, dtype=np.complex128)[:, 0], 1, alpha, beta
            ene, z = _tridiag_lowest(np.array(alpha), np.array(beta))
            conv = abs(tmp / ene - 1.0)
            if i == maxitr or conv < conv_tol:
                break
            tmp = ene
    if not want_vector:
        return ene, None, ite, alpha, beta
    psi1 = psi_io.copy()
    psi_io = psi_io * z[0]
    psi3 = np.zeros_like(psi1)
    for i in range(1, ite):
        psi2 = hv(psi1)
        if i > 1:
            psi2 = psi2 - alpha[i - 1] * psi1 - beta[i - 2] * psi3
        else:
            psi2 = psi2 - alpha[i - 1] * psi1
        psi3 = psi1
        psi1 = psi2 / beta[i - 1]
        psi_io = psi_io + z[i] * psi1
    psi_io /= np.linalg.norm(psi_io)
    return ene, psi_io, ite, alpha, beta


def trlanczos(hv, v0, NOK, NOM, maxitr=1000, conv_tol=1e-14, NOE=None):
    """lancz_eigen_val_complex (eigen_solver.f90:69-261): thick-restart Lanczos
    with full reorthogonalisation against psi_1..psi_{i-1}.  Returns
    (ene[NOM], psi[:, :NOK] Ritz vectors, restarts)."""
    if NOM < 2 * NOK:
        NOM = 2 * NOK
    dim = len(v0)
    psi = np.zeros((dim, NOM + 1), dtype=np.complex128)
    psi[:, 0] = v0
    alpha = np.zeros(NOM)
    beta = np.zeros(NOM + 1)
    beta[0] = np.linalg.norm(psi[:, 0])

    def reorth(i):  # full_reorthogonalization_complex, j = 1..i-1 (1-based)
        for j in range(i - 1):
            ov = np.vdot(psi[:, j], psi[:, i])
            psi[:, i] -= ov * psi[:, j]

    for i in range(1, NOM + 1):
        psi[:, i - 1] /= beta[i - 1]
        Q = hv(psi[:, i - 1])
        alpha[i - 1] = np.vdot(psi[:, i - 1], Q).real
        if i > 1:
            psi[:, i] = Q - alpha[i - 1] * psi[:, i - 1] - beta[i - 1] * psi[:, i - 2]
            reorth(i)
        else:
            psi[:, i] = Q - alpha[i - 1] * psi[:, i - 1]
        beta[i] = np.linalg.norm(psi[:, i])
    psi[:, NOM] /= beta[NOM]
    T = np.diag(alpha) + np.diag(beta[1:NOM], 1) + np.diag(beta[1:NOM], -1)
    ene, coef = np.linalg.eigh(T)
    ccoef = coef.astype(np.complex128)
    calpha = alpha.astype(np.complex128)
    cbeta = beta.astype(np.complex128)
    flag = 0
    tritr = 0
    for tritr in range(1, maxitr + 1):
        psi[:, :NOK] = psi[:, :NOM] @ ccoef[:, :NOK]
        conv = np.abs(calpha[:NOK] / ene[:NOK] - 1.0)
        flag = flag + 1 if conv.max() < conv_tol else 0
        if flag == 2 or tritr == maxitr:
            break
        psi[:, NOK] = psi[:, NOM]
        calpha[:NOK] = ene[:NOK]
        cb = cbeta[NOM] * ccoef[NOM - 1, :NOK]
        cbeta[1:NOK + 1] = cb
        i = NOK + 1
        Q = hv(psi[:, i - 1])
        calpha[i - 1] = np.vdot(psi[:, i - 1], Q)
        for j in range(NOK):
            Q = Q - cbeta[j + 1] * psi[:, j]
        psi[:, i] = Q - calpha[i - 1] * psi[:, i - 1]
        reorth(i)
        cbeta[i] = np.linalg.norm(psi[:, i])
        for i in range(NOK + 2, NOM + 1):
            psi[:, i - 1] /= cbeta[i - 1]
            Q = hv(psi[:, i - 1])
            calpha[i - 1] = np.vdot(psi[:, i - 1], Q)
            psi[:, i] = Q - calpha[i - 1] * psi[:, i - 1] - cbeta[i - 1] * psi[:, i - 2]
            reorth(i)
            cbeta[i] = np.linalg.norm(psi[:, i])
        psi[:, NOM] /= cbeta[NOM]
        T = np.zeros((NOM, NOM), dtype=np.complex128)
        for j in range(NOM):
            T[j, j] = calpha[j]
        # upper triangle as in eigen_solver.f90:196-203: T(1:NOK,NOK+1) = cbeta(1:NOK),
        # T(j,j+1) = cbeta(j) for j = NOK+1..NOM-1 (1-based)
        T[:NOK, NOK] = cbeta[1:NOK + 1]
        for j1 in range(NOK + 1, NOM):
            T[j1 - 1, j1] = cbeta[j1]
        ene, ccoef = np.linalg.eigh(T, UPLO="U")
    return ene, psi[:, :NOK].copy(), tritr


def dsf_recurrence(hv, f, itr_dsf, rfield=0.0, sztot_new=0.0):
    """calcu_DSF_wavevector recurrence (get_expectation_values.f90:481-592):
    returns alpha[itr], beta2[itr] (beta2[0] = 0, beta2[n] = ||f_{n+1}'||^2) and
    norm1 = <f|f>, with f = S^a_q |psi> supplied."""
    dim = len(f)
    norm1 = float(np.vdot(f, f).real)
    fn = f / math.sqrt(norm1)
    alpha = np.full(itr_dsf, 1.0e300)
    beta2 = np.zeros(itr_dsf + 1)
    fnp1 = hv(fn) - rfield * sztot_new * fn
    alpha[0] = float(np.vdot(fn, fnp1).real)
    if dim > 1:
        fnp1 = fnp1 - alpha[0] * fn
        beta2[1] = float(np.vdot(fnp1, fnp1).real)
        fnp1 = fnp1 / math.sqrt(beta2[1])
        for n in range(2, itr_dsf + 1):
            fnm1, fn = fn, fnp1
            fnp1 = hv(fn) - rfield * sztot_new * fn
            alpha[n - 1] = float(np.vdot(fn, fnp1).real)
            if n == dim:
                break
            fnp1 = fnp1 - alpha[n - 1] * fn - math.sqrt(beta2[n - 1]) * fnm1
            beta2[n] = float(np.vdot(fnp1, fnp1).real)
            fnp1 = fnp1 / math.sqrt(beta2[n])
    return alpha, beta2[:itr_dsf], norm1


def continued_fraction(alpha, beta2, norm1, E0, omegas, eta, reference_quirk=True):
    """DSF_analysis (source_DSFLan/main.f90:52-84): S(w) = -Im G / pi with
    G = norm1 / (z - a1 - b2/(z - a2 - ...)), z = w + E0 + i eta.  beta2[n-1]
    holds the file's beta(n) (beta(1) = 0).  reference_quirk=True reproduces the
    deepest level exactly as DSFLan evaluates it (imaginary part not divided by
    the denominator, main.f90:64-65); False gives the exact continued fraction."""
    a = np.asarray(alpha, dtype=np.float64)
    b = np.concatenate([np.asarray(beta2, dtype=np.float64), [0.0]])  # b[n-1] = beta(n), beta(M+1) = 0
    M = len(a)
    w = np.asarray(omegas, dtype=np.float64)
    kr = w + E0 - a[M - 1] - b[M]
    bunbo = kr * kr + eta * eta
    tr = b[M - 1] * kr / bunbo
    ti = -b[M - 1] * eta if reference_quirk else -b[M - 1] * eta / bunbo
    for n in range(M - 1, 1, -1):
        kr = w + E0 - a[n - 1] - tr
        ki = eta - ti
        bunbo = kr * kr + ki * ki
        tr = b[n - 1] * kr / bunbo
        ti = -b[n - 1] * ki / bunbo
    kr = w + E0 - a[0] - tr
    ki = eta - ti
    bunbo = kr * kr + ki * ki
    ti = -norm1 * ki / bunbo
    return -ti / math.acos(-1.0)


# ---------------------------------------------------------------------------
# The reference's own modules (oracle/_ref/libqs3ref.so)
# ---------------------------------------------------------------------------
class RefLibU1:
    """ctypes wrapper around oracle/ref_harness_u1.f90: the reference's own
    U(1)-only modules (source_only_u1/).  One model at a time."""

    def __init__(self):
        if not os.path.exists(_REF_U1_PATH):
            raise FileNotFoundError(_REF_U1_PATH)
        self.L = ctypes.CDLL(_REF_U1_PATH)
        self.THS = 0

    @staticmethod
    def available():
        return os.path.exists(_REF_U1_PATH)

    def setup(self, NOS, nod, bonds):
        b = np.asarray(bonds, dtype=np.float64).reshape(-1, 4)
        self._keep = [np.ascontiguousarray(b[:, c].astype(t)) for c, t in
                      ((0, np.int64), (1, np.int64), (2, np.float64), (3, np.float64))]
        pi, pj, jxy, jz = self._keep
        t = np.zeros(1, dtype=np.int64)
        self.L.ref_u1_setup(i64(NOS), i64(nod), i64(len(b)), _ptr(pi), _ptr(pj), _ptr(jxy), _ptr(jz), _ptr(t))
        self.THS = int(t[0])
        return self.THS

    def hv(self, v):
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1)
        self.L.ref_u1_hv(_ptr(v0), _ptr(v1))
        return v0


class RefLib:
    """Thin ctypes wrapper around oracle/ref_harness.f90.  The Fortran modules
    keep global state: one model / sector at a time."""

    def __init__(self):
        if not os.path.exists(_REF_PATH):
            raise FileNotFoundError(_REF_PATH)
        self.L = ctypes.CDLL(_REF_PATH)
        self.THS = 0
        self.THS_new = 0

    @staticmethod
    def available():
        return os.path.exists(_REF_PATH)

    def setup(self, NOS, LX, LY, LZ, nod, K, bonds):
        b = np.asarray(bonds, dtype=np.float64).reshape(-1, 4)
        self._keep = [np.ascontiguousarray(b[:, c].astype(t)) for c, t in
                      ((0, np.int64), (1, np.int64), (2, np.float64), (3, np.float64))]
        pi, pj, jxy, jz = self._keep
        self.L.ref_setup(i64(NOS), i64(LX), i64(LY), i64(LZ), i64(nod), i64(K[0]), i64(K[1]),
                         i64(K[2]), i64(len(b)), _ptr(pi), _ptr(pj), _ptr(jxy), _ptr(jz))
        self.NOS, self.LX, self.LY, self.LZ, self.nod = NOS, LX, LY, LZ, nod

    def build_sector(self):
        t = np.zeros(1, dtype=np.int64)
        self.L.ref_build_sector(_ptr(t))
        self.THS = int(t[0])
        return self.THS

    def lists(self):
        ls = np.zeros(self.THS, dtype=np.int64)
        lr = np.zeros(self.THS, dtype=np.float64)
        self.L.ref_get_lists(_ptr(ls), _ptr(lr))
        return ls, lr

    def explist(self):
        ex = np.zeros((self.LX + 1) * (self.LY + 1) * (self.LZ + 1), dtype=np.complex128)
        self.L.ref_explist(_ptr(ex))
        return ex

    def hv(self, v):
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1)
        self.L.ref_hv(_ptr(v0), _ptr(v1))
        return v0

    def set_lists(self, list_s, list_r):
        """Install lists built elsewhere as the module's list_s / list_r."""
        self._lists = (np.ascontiguousarray(list_s, dtype=np.int64), np.ascontiguousarray(list_r, dtype=np.float64))
        self.L.ref_set_lists(i64(len(self._lists[0])), _ptr(self._lists[0]), _ptr(self._lists[1]))
        self.THS = len(self._lists[0])

    def hv_rows(self, v, r0, r1, out=None):
        """Rows r0..r1-1 (0-based) of ham_to_vec_wave_vector's loop
        (ham2vec_v3.f90:52-70) with the reference's own routines; the other
        rows of the returned vector are left as they were in out."""
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1) if out is None else out
        self.L.ref_hv_rows(i64(r0 + 1), i64(r1), _ptr(v0), _ptr(v1))
        return v0

    def representative(self, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.int64)).copy()
        t = np.zeros(1, dtype=np.int64)
        ell = np.zeros(3, dtype=np.int64)
        self.L.ref_representative(i64(len(a)), _ptr(a), _ptr(t), _ptr(ell))
        return int(t[0]), ell, a

    def checkstate(self, s, nod):
        ro = np.zeros(1)
        self.L.ref_checkstate(i64(s), i64(nod), _ptr(ro))
        return float(ro[0])

    def unrank(self, t, nod):
        out = np.zeros(nod, dtype=np.int64)
        self.L.ref_list_fly(i64(t), i64(nod), _ptr(out))
        return out

    def j_flip(self, i, j, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.int64))
        out = np.zeros(len(a), dtype=np.int64)
        self.L.ref_j_flip(i64(i), i64(j), i64(len(a)), _ptr(a), _ptr(out))
        return out

    def cf_pair(self, r, v):
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        no = np.zeros(1, dtype=np.int64)
        sz = np.zeros(1)
        sp = np.zeros(1, dtype=np.complex128)
        self.L.ref_cf_pair(i64(r[0]), i64(r[1]), _ptr(v1), _ptr(no), _ptr(sz), _ptr(sp))
        return int(no[0]), float(sz[0]), complex(sp[0])

    def lm_site(self, n, v):
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        no = np.zeros(1, dtype=np.int64)
        val = np.zeros(1)
        self.L.ref_lm_site(i64(n), _ptr(v1), _ptr(no), _ptr(val))
        return int(no[0]), float(val[0])

    def build_new_sector(self, spsmsz, q):
        t = np.zeros(1, dtype=np.int64)
        self.L.ref_build_new_sector(i64(spsmsz), dbl(q[0]), dbl(q[1]), dbl(q[2]), _ptr(t))
        self.THS_new = int(t[0])
        return self.THS_new

    def new_lists(self):
        ls = np.zeros(self.THS_new, dtype=np.int64)
        lr = np.zeros(self.THS_new, dtype=np.float64)
        self.L.ref_get_new_lists(_ptr(ls), _ptr(lr))
        return ls, lr

    def sq(self, pos, psi):
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        v = np.ascontiguousarray(psi, dtype=np.complex128)
        fn = np.zeros(self.THS_new, dtype=np.complex128)
        self.L.ref_sq(_ptr(p), _ptr(v), _ptr(fn))
        return fn

    def hv_new(self, v):
        v1 = np.ascontiguousarray(v, dtype=np.complex128)
        v0 = np.zeros_like(v1)
        self.L.ref_hv_new(_ptr(v0), _ptr(v1))
        return v0
