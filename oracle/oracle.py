"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (kpm_oracle.c, the FP64 restatement of RSATB's KPM
path) and adds the dense numpy references the tests pin it with. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product never does.

Dense references (numpy LAPACK, our own matrices):
  - dense_moments: mu_n = sum_k |<psi_k|r>|^2 T_n(eps_k) from eigh of the
    scaled operator, the exact value of the stochastic estimator for the
    same r (rsatb/calc/dos.hpp:48-60 defines the estimator).
  - isqrt_patches: per block row i, A = S[C_i, C_i] and Z[i, C_i] = the rows
    of A^{-1/2} belonging to block i, with f(x) = x^{-1/2} for x > 1e-12 else
    0 -- the clamped eigendecomposition route of rsatb/calc/isqrtm.hpp:167-173
    over the subgraph problems of graph_inverse_sqrt (isqrtm.hpp:137-186).
  - dense_group_moments: sum over a group's rows of diag T_m(A) by the dense
    recurrence, the check rsatb/calc/test/test_pdos.cpp:196-229 makes.
  - dos_grid / compute_dos: ComputeDos's kpm branch
    (rsatb/calc/observables.hpp:336-350, 370-393, 439-443).
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

FORM_ORTHOGONAL, FORM_SANDWICH = 0, 1


class _Bsr(C.Structure):
    _fields_ = [
        ("nbr", C.c_int64),
        ("b", C.c_int32),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col_ind", C.POINTER(C.c_int32)),
        ("vals", C.POINTER(C.c_double)),
    ]


class _Op(C.Structure):
    _fields_ = [
        ("form", C.c_int32),
        ("H", C.POINTER(_Bsr)),
        ("Z", C.POINTER(_Bsr)),
        ("center", C.c_double),
        ("half_width", C.c_double),
        ("ZH", C.POINTER(_Bsr)),
    ]


_lib = None


def build() -> Path:
    import subprocess

    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    pd = C.POINTER(C.c_double)
    lib.or_mix64.restype = C.c_uint64
    lib.or_mix64.argtypes = [C.c_uint64]
    lib.or_rademacher_sign.restype = C.c_double
    lib.or_rademacher_sign.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.or_samples_rademacher.argtypes = [C.c_uint64, C.c_int64, C.c_int32, pd]
    lib.or_samples_phase.argtypes = [C.c_uint64, C.c_int64, C.c_int32, pd]
    lib.or_bsr_mult.argtypes = [C.POINTER(_Bsr), pd, C.c_int32, pd]
    lib.or_bsr_mult_adjoint.argtypes = [C.POINTER(_Bsr), pd, C.c_int32, pd]
    lib.or_apply_op.argtypes = [C.POINTER(_Op), pd, C.c_int32, pd, pd]
    lib.or_apply_op.restype = C.c_int
    lib.or_kpm_moments.argtypes = [C.POINTER(_Op), pd, C.c_int32, C.c_int32, C.c_int32, C.c_int32, pd, pd]
    lib.or_kpm_moments.restype = C.c_int
    lib.or_kpm_pdos_moments.argtypes = [C.POINTER(_Op), pd, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        C.POINTER(C.c_int32), pd, pd]
    lib.or_kpm_pdos_moments.restype = C.c_int
    lib.or_jackson_kernel.argtypes = [C.c_int32, pd]
    lib.or_kpm_kernel_weights.argtypes = [C.c_int32, C.c_int32, pd]
    lib.or_kpm_kernel_weights.restype = C.c_int
    lib.or_kpm_density.argtypes = [pd, C.c_int32, pd, C.c_int64, C.c_double, C.c_int32, pd]
    lib.or_kpm_density.restype = C.c_int
    lib.or_num_threads.restype = C.c_int
    lib.or_set_num_threads.argtypes = [C.c_int32]
    _lib = lib
    return lib


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _Holder:
    """Keeps the numpy arrays behind a ctypes struct alive."""

    def __init__(self, bsr):
        self.rp = np.ascontiguousarray(bsr.row_ptr, dtype=np.int64)
        self.ci = np.ascontiguousarray(bsr.col_ind, dtype=np.int32)
        self.v = np.ascontiguousarray(np.asarray(bsr.values, dtype=np.complex128))
        self.s = _Bsr(bsr.n_block_rows, bsr.block_dim, self.rp.ctypes.data_as(C.POINTER(C.c_int64)),
                      self.ci.ctypes.data_as(C.POINTER(C.c_int32)), _pd(self.v.view(np.float64)))


def _cvec(x) -> np.ndarray:
    x = np.asarray(x)
    if x.ndim == 1:
        x = x[:, None]
    return np.ascontiguousarray(x, dtype=np.complex128)


# ------------------------------------------------------------------- RNG
def mix64(x: int) -> int:
    return int(load().or_mix64(C.c_uint64(x)))


def rademacher_sign(seed: int, row: int, state: int) -> float:
    return float(load().or_rademacher_sign(seed, row, state))


def samples_rademacher(seed: int, rows: int, R: int) -> np.ndarray:
    out = np.zeros((rows, R), dtype=np.complex128)
    load().or_samples_rademacher(seed, rows, R, _pd(out.view(np.float64)))
    return out


def samples_phase(seed: int, rows: int, R: int) -> np.ndarray:
    out = np.zeros((rows, R), dtype=np.complex128)
    load().or_samples_phase(seed, rows, R, _pd(out.view(np.float64)))
    return out


# ------------------------------------------------------------ operators
def bsr_mult(A, x) -> np.ndarray:
    h = _Holder(A)
    xc = _cvec(x)
    y = np.zeros_like(xc)
    load().or_bsr_mult(C.byref(h.s), _pd(xc.view(np.float64)), xc.shape[1], _pd(y.view(np.float64)))
    return y


def bsr_mult_adjoint(A, x) -> np.ndarray:
    h = _Holder(A)
    xc = _cvec(x)
    y = np.zeros_like(xc)
    load().or_bsr_mult_adjoint(C.byref(h.s), _pd(xc.view(np.float64)), xc.shape[1], _pd(y.view(np.float64)))
    return y


class Operator:
    """The scaled operator (A - center)/half_width, A = H or Z^H H Z."""

    def __init__(self, form: int, H, center: float, half_width: float, Z=None, ZH=None):
        """ZH (optional): an explicit Z^H applied by the row-gather kernel
        instead of scattering Z's blocks (the faster CPU variant)."""
        self.form = form
        self._h = _Holder(H)
        self._z = _Holder(Z) if Z is not None else None
        self._zh = _Holder(ZH) if ZH is not None else None
        self.rows = H.n_block_rows * H.block_dim
        self.op = _Op(form, C.pointer(self._h.s), C.pointer(self._z.s) if self._z else None, center, half_width,
                      C.pointer(self._zh.s) if self._zh else None)

    def apply(self, x) -> np.ndarray:
        xc = _cvec(x)
        y = np.zeros_like(xc)
        tmp = np.zeros_like(xc)
        rc = load().or_apply_op(C.byref(self.op), _pd(xc.view(np.float64)), xc.shape[1],
                                _pd(y.view(np.float64)), _pd(tmp.view(np.float64)))
        if rc != 0:
            raise RuntimeError("or_apply_op failed")
        return y

    def moments(self, psi, M: int, doubling: bool = True, fused: bool = False, per_column: bool = False):
        pc = _cvec(psi)
        R = pc.shape[1]
        mu = np.zeros(M, dtype=np.complex128)
        cols = np.zeros((M, R), dtype=np.complex128)
        rc = load().or_kpm_moments(C.byref(self.op), _pd(pc.view(np.float64)), R, M, int(doubling), int(fused),
                                   _pd(mu.view(np.float64)), _pd(cols.view(np.float64)))
        if rc != 0:
            raise RuntimeError("or_kpm_moments failed")
        return (mu, cols) if per_column else mu

    def pdos_moments(self, psi, M: int, group_of_row, n_groups: int, doubling: bool = True):
        """kpm_pdos_moments: (total [M], group [n_groups, M]) complex."""
        pc = _cvec(psi)
        R = pc.shape[1]
        grp = np.ascontiguousarray(np.asarray(group_of_row, dtype=np.int32))
        if grp.size != self.rows:
            raise RuntimeError("pdos: group_of_row must map every scalar row")
        total = np.zeros(M, dtype=np.complex128)
        group = np.zeros((n_groups, M), dtype=np.complex128)
        rc = load().or_kpm_pdos_moments(C.byref(self.op), _pd(pc.view(np.float64)), R, M, int(doubling),
                                        n_groups, grp.ctypes.data_as(C.POINTER(C.c_int32)),
                                        _pd(total.view(np.float64)), _pd(group.view(np.float64)))
        if rc == -2:
            raise RuntimeError("pdos: group id exceeds n_groups")
        if rc != 0:
            raise RuntimeError("or_kpm_pdos_moments failed")
        return total, group


# ---------------------------------------------------------------- KPM
def jackson_kernel(n: int) -> np.ndarray:
    out = np.zeros(n)
    load().or_jackson_kernel(n, _pd(out))
    return out


def kernel_weights(kernel: int, n: int) -> np.ndarray:
    g = np.zeros(n)
    if load().or_kpm_kernel_weights(kernel, n, _pd(g)) != 0:
        raise RuntimeError("bad kernel")
    return g


def kpm_density(mu, x, edge_inset: float, kernel: int = 1) -> np.ndarray:
    m = np.ascontiguousarray(np.asarray(mu, dtype=np.complex128))
    xv = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    rho = np.zeros(xv.size)
    load().or_kpm_density(_pd(m.view(np.float64)), m.size, _pd(xv), xv.size, edge_inset, kernel, _pd(rho))
    return rho


def dos_plan(emin: float, emax: float, de: float):
    span = emax - emin
    half = span / 2.0
    nt = max(1, int(math.ceil(span / de)))
    n = max(1, int(math.ceil(math.sqrt(8.0 * math.log(2.0)) * math.pi * half / de)))
    return nt, n


def dos_grid(nt: int) -> np.ndarray:
    x = np.zeros(2 * nt)
    for i in range(nt):
        x[i] = (i - nt) / nt
        x[nt + i] = i / nt
    return x


def compute_dos(op: Operator, psi, emin: float, emax: float, de: float):
    """ComputeDos kpm branch on one k-point: (dos, omega)."""
    nt, n = dos_plan(emin, emax, de)
    mu = op.moments(psi, n, doubling=True)
    rho = kpm_density(mu, dos_grid(nt), 0.5 / nt, 1)
    half, shift = (emax - emin) / 2.0, (emax + emin) / 2.0
    omega = np.array([(i - nt) * half / nt + shift for i in range(nt)] + [i * half / nt + shift for i in range(nt)])
    return 2.0 * rho, omega


def set_threads(n: int) -> None:
    load().or_set_num_threads(n)


def num_threads() -> int:
    return int(load().or_num_threads())


# ------------------------------------------------------- dense references
def dense_scaled(form: int, H, center: float, half_width: float, Z=None) -> np.ndarray:
    Hd = H.to_dense().astype(np.complex128)
    if form == FORM_SANDWICH:
        Zd = Z.to_dense().astype(np.complex128)
        A = Zd.conj().T @ Hd @ Zd
    else:
        A = Hd
    n = A.shape[0]
    return (A - center * np.eye(n)) / half_width


def dense_group_moments(A: np.ndarray, group_of_row, n_groups: int, M: int) -> np.ndarray:
    """ref[g, m] = sum_{i in g} [T_m(A)]_ii by the plain three-term recurrence
    on dense matrices (the check of rsatb/calc/test/test_pdos.cpp:196-229);
    row n_groups is the all-rows channel."""
    n = A.shape[0]
    grp = np.asarray(group_of_row)
    ref = np.zeros((n_groups + 1, M))
    T0, T1 = np.eye(n, dtype=A.dtype), A.copy()

    def record(m, T):
        d = np.real(np.diag(T))
        ref[n_groups, m] = d.sum()
        mask = grp >= 0
        np.add.at(ref[:, m], grp[mask], d[mask])

    record(0, T0)
    if M > 1:
        record(1, T1)
    for m in range(2, M):
        T2 = 2.0 * (A @ T1) - T0
        record(m, T2)
        T0, T1 = T1, T2
    return ref


def dense_moments(A: np.ndarray, psi, M: int) -> np.ndarray:
    """Column-mean moments <r|T_n(A)|r> via eigh of Hermitian A."""
    Ah = 0.5 * (A + A.conj().T)
    w, V = np.linalg.eigh(Ah)
    pc = _cvec(psi)
    wt = np.abs(V.conj().T @ pc) ** 2  # [n_eig, R]
    th = np.arccos(np.clip(w, -1.0, 1.0))
    T = np.cos(np.outer(np.arange(M), th))  # T_n(w) = cos(n arccos w) on [-1, 1]
    return (T @ wt).mean(axis=1).astype(np.complex128)


class Bsr:
    """Plain BSR holder with the attribute names the library binding reads."""

    def __init__(self, n_block_rows, block_dim, row_ptr, col_ind, values):
        self.n_block_rows, self.block_dim = int(n_block_rows), int(block_dim)
        self.row_ptr, self.col_ind, self.values = row_ptr, col_ind, values

    def to_dense(self) -> np.ndarray:
        b = self.block_dim
        n = self.n_block_rows * b
        out = np.zeros((n, n), dtype=self.values.dtype)
        for i in range(self.n_block_rows):
            for s in range(int(self.row_ptr[i]), int(self.row_ptr[i + 1])):
                j = int(self.col_ind[s])
                out[i * b:(i + 1) * b, j * b:(j + 1) * b] += self.values[s]
        return out


def isqrt_patches(S) -> Bsr:
    """Seed Z of graph_inverse_sqrt by the clamped eigen route (BSR like S)."""
    b = S.block_dim
    out = np.zeros_like(S.values)
    for i in range(S.n_block_rows):
        s0, s1 = int(S.row_ptr[i]), int(S.row_ptr[i + 1])
        ci = S.col_ind[s0:s1].astype(np.int64)
        n = ci.size * b
        A = np.zeros((n, n), dtype=S.values.dtype)
        pos = {int(c): k for k, c in enumerate(ci)}
        for k, c in enumerate(ci):
            r0, r1 = int(S.row_ptr[c]), int(S.row_ptr[c + 1])
            for s in range(r0, r1):
                cc = int(S.col_ind[s])
                if cc in pos:
                    kk = pos[cc]
                    A[k * b:(k + 1) * b, kk * b:(kk + 1) * b] = S.values[s]
        w, V = np.linalg.eigh(0.5 * (A + A.conj().T))
        f = np.where(w > 1e-12, 1.0 / np.sqrt(np.maximum(w, 1e-300)), 0.0)
        Ainv = (V * f) @ V.conj().T
        me = pos[i]
        rowblk = Ainv[me * b:(me + 1) * b, :]
        for k in range(ci.size):
            out[s0 + k] = rowblk[:, k * b:(k + 1) * b]
    return Bsr(S.n_block_rows, b, S.row_ptr.copy(), S.col_ind.copy(), out)


# ------------------------------------------------- dense moment gradient
def chebyshev_divided_differences(lam: np.ndarray, m: int) -> np.ndarray:
    """D[a, b] = (T_m(l_a) - T_m(l_b)) / (l_a - l_b), and T_m'(l_a) on the
    diagonal, from l = cos(t): a ratio of sines that stays accurate for close
    eigenvalues."""
    th = np.arccos(np.clip(lam, -1.0, 1.0))
    s = 0.5 * (th[:, None] + th[None, :])
    d = 0.5 * (th[:, None] - th[None, :])
    num = np.sin(m * s) * np.sin(m * d)
    den = np.sin(s) * np.sin(d)
    diag_like = np.abs(d) < 1e-7
    with np.errstate(divide="ignore", invalid="ignore"):
        D = num / den
    # limit d -> 0: m sin(m s) / sin(s) = T_m'(cos s)
    lim = m * np.sin(m * s) / np.sin(s)
    return np.where(diag_like, lim, D)


def dense_moment_gradient(A: np.ndarray, psi, g: np.ndarray) -> np.ndarray:
    """dL/dA for L = sum_m g_m Re mu_m, mu_m = mean_c <r_c|T_m(A)|r_c>, A real
    symmetric with spectrum in (-1, 1): the Daleckii-Krein formula
    G = V (F o W) V^T, W_ab = mean_c Re(conj(w_a) w_b), w = V^T r,
    F = sum_m g_m D_m (divided differences of T_m)."""
    w, V = np.linalg.eigh(0.5 * (A + A.T))
    pc = _cvec(psi)
    Wc = V.T @ pc
    Wt = np.real(np.conj(Wc) @ Wc.T) / pc.shape[1]
    F = np.zeros_like(Wt)
    for m, gm in enumerate(g):
        if gm != 0.0 and m > 0:
            F += gm * chebyshev_divided_differences(w, m)
    return V @ (F * Wt) @ V.T


# ------------------------------------------ gradient through the patches
def _patch_of(S, i):
    """(block columns C_i, dense A_i = S[C_i, C_i], position of block i)."""
    b = S.block_dim
    s0, s1 = int(S.row_ptr[i]), int(S.row_ptr[i + 1])
    ci = S.col_ind[s0:s1].astype(np.int64)
    pos = {int(c): k for k, c in enumerate(ci)}
    A = np.zeros((ci.size * b, ci.size * b))
    for k, c in enumerate(ci):
        for s in range(int(S.row_ptr[c]), int(S.row_ptr[c + 1])):
            cc = int(S.col_ind[s])
            if cc in pos:
                kk = pos[cc]
                A[k * b:(k + 1) * b, kk * b:(kk + 1) * b] = S.values[s]
    return ci, pos, A


def isqrt_patches_vjp(S, dZ: np.ndarray) -> np.ndarray:
    """dL/dS on S's pattern ([nnzb, b, b]) for dL/dZ on Z's (= S's) pattern,
    through isqrt_patches (real S). Per patch, A = S[C_i, C_i] = V diag(w) V^T
    and X = A^{-1/2}: dL/dA = V (F o (V^T E V)) V^T (Daleckii-Krein), with E =
    dL/dX (dL/dZ's block row i in the rows of block i) and F the divided
    differences of f(x) = x^{-1/2}, in the cancellation-free form
    (f(a) - f(b)) / (a - b) = -1 / (sqrt(a) sqrt(b) (sqrt(a) + sqrt(b)))
    (= f'(a) at a = b). Each entry of dL/dA is added to the stored block of S
    it was gathered from."""
    b = S.block_dim
    out = np.zeros_like(S.values, dtype=np.float64)
    for i in range(S.n_block_rows):
        s0 = int(S.row_ptr[i])
        ci, pos, A = _patch_of(S, i)
        n = A.shape[0]
        E = np.zeros((n, n))
        me = pos[i]
        for k in range(ci.size):
            E[me * b:(me + 1) * b, k * b:(k + 1) * b] = dZ[s0 + k]
        w, V = np.linalg.eigh(0.5 * (A + A.T))
        sq = np.sqrt(w)
        F = -1.0 / (sq[:, None] * sq[None, :] * (sq[:, None] + sq[None, :]))
        dA = V @ (F * (V.T @ E @ V)) @ V.T
        for k, c in enumerate(ci):
            for s in range(int(S.row_ptr[c]), int(S.row_ptr[c + 1])):
                cc = int(S.col_ind[s])
                if cc in pos:
                    kk = pos[cc]
                    out[s] += dA[k * b:(k + 1) * b, kk * b:(kk + 1) * b]
    return out


def _on_pattern(P, D: np.ndarray) -> np.ndarray:
    """Entries of the dense matrix D at the stored blocks of P ([nnzb, b, b])."""
    b = P.block_dim
    out = np.zeros((int(P.row_ptr[-1]), b, b))
    for i in range(P.n_block_rows):
        for s in range(int(P.row_ptr[i]), int(P.row_ptr[i + 1])):
            j = int(P.col_ind[s])
            out[s] = D[i * b:(i + 1) * b, j * b:(j + 1) * b]
    return out


def sandwich_moment_gradients(H, S, center: float, half_width: float, psi, g: np.ndarray) -> dict:
    """Gradients of L = sum_m g_m Re mu_m for the sandwich operator
    A~ = (Z^T H Z - center) / half_width with Z = isqrt_patches(S), from the
    dense eigendecomposition: G = dL/dA~ (dense_moment_gradient, symmetric),
    dL/dH = Z G Z^T / a, dL/dZ = (H Z G + H Z G^T) / a, dL/dS =
    isqrt_patches_vjp(S, dL/dZ); each on its matrix's pattern, plus the
    moments themselves. Real H and S."""
    Z = isqrt_patches(S)
    Zd, Hd = Z.to_dense(), H.to_dense()
    n = Hd.shape[0]
    A = (Zd.T @ Hd @ Zd - center * np.eye(n)) / half_width
    G = dense_moment_gradient(A, psi, g)
    dH = Zd @ G @ Zd.T / half_width
    dZ = (Hd @ Zd @ G + Hd @ Zd @ G.T) / half_width
    dZp = _on_pattern(Z, dZ)
    return {"H": _on_pattern(H, dH), "Z": dZp, "S": isqrt_patches_vjp(S, dZp),
            "mu": dense_moments(A, psi, len(g)), "Z_bsr": Z}
