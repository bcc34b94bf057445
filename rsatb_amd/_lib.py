"""ctypes binding of librsatb_gpu.so (the C ABI of include/rsatb_gpu.h).

The library is the only compute path: there is no CPU fallback. Loading
fails loudly when the shared object has not been built, and every call that
needs a device fails loudly when there is none.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "librsatb_gpu.so"

RG_OK = 0
RG_ERR_INVALID = 1
RG_ERR_HIP = 2
RG_ERR_STATE = 3
RG_ERR_CANCELLED = 4
RG_ERR_NOMEM = 5
RG_ERR_UNSUPPORTED = 6

SLOT_H, SLOT_Z, SLOT_ZH, SLOT_S = 0, 1, 2, 3
SLOT_VX, SLOT_VY, SLOT_VZ = 4, 5, 6
SLOT_DH, SLOT_DVX, SLOT_DVY, SLOT_DVZ = 7, 8, 9, 10
F64, F32, C128, C64 = 0, 1, 2, 3
FORM_ORTHOGONAL, FORM_SANDWICH = 0, 1
SAMPLES_RADEMACHER, SAMPLES_PHASE, SAMPLES_UPLOAD_REAL, SAMPLES_UPLOAD_COMPLEX = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
DOUBLING_AUTO, DOUBLING_OFF, DOUBLING_ON = -1, 0, 1
KERNEL_NONE, KERNEL_JACKSON, KERNEL_COSTAPER, KERNEL_VPTAPER = 0, 1, 2, 3
KCLASS_SPMM_PLAIN, KCLASS_SPMM_STEP, KCLASS_REDUCE, KCLASS_GROUP, KCLASS_KG_GEMM = 0, 1, 2, 3, 4
KCLASS_NAMES = {KCLASS_SPMM_PLAIN: "spmm_plain", KCLASS_SPMM_STEP: "spmm_step", KCLASS_REDUCE: "reduce",
                KCLASS_GROUP: "group_dots", KCLASS_KG_GEMM: "kg_gemm"}

# Every symbol include/rsatb_gpu.h declares, with (restype, argtypes).
_vp, _i, _i64, _d, _u64 = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_uint64
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)


class SkTableC(C.Structure):
    """struct rg_sk_table (include/rsatb_gpu.h)."""

    _fields_ = [
        ("n_types", C.c_int32), ("norb", C.c_int32), ("n_channels", C.c_int32),
        ("orb_l", _pi32), ("orb_m", _pi32), ("onsite", _pd), ("pair_cutoff", _pd),
        ("ch_row_type", _pi32), ("ch_col_type", _pi32), ("ch_row_l", _pi32), ("ch_col_l", _pi32),
        ("ch_bond", _pi32), ("ch_kind", _pi32), ("ch_off", _pi32), ("ch_n", _pi32),
        ("ch_r_cut", _pd), ("ch_lambda", _pd), ("ch_r_on", _pd),
        ("rad_x", _pd), ("rad_y", _pd), ("rad_y2", _pd), ("n_rad", C.c_int64),
    ]


SIGNATURES = {
    "rg_version": (_i, []),
    "rg_create": (_i, [_i, C.POINTER(_vp)]),
    "rg_destroy": (None, [_vp]),
    "rg_last_error": (C.c_char_p, [_vp]),
    "rg_mem_info": (_i, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "rg_upload_bsr": (_i, [_vp, _i, _i64, _i, _pi64, _pi32, _vp, _i]),
    "rg_set_operator": (_i, [_vp, _i, _d, _d]),
    "rg_set_samples": (_i, [_vp, _i64, _i, _i, _u64, _pd]),
    "rg_get_samples": (_i, [_vp, _pd, _i]),
    "rg_apply_op": (_i, [_vp, _i, _pd, _pd, _i]),
    "rg_bsr_mult": (_i, [_vp, _i, _i, _pd, _pd, _i]),
    "rg_kpm_moments": (_i, [_vp, _i, _i, _i, _pd, _pd]),
    "rg_kpm_moments_noalloc": (_i, [_vp, _i, _i, _i, _pd, _pd]),
    "rg_kpm_pdos_moments": (_i, [_vp, _i, _i, _i, _i64, _pi32, _i, _pd, _pd]),
    "rg_dos_plan": (_i, [_d, _d, _d, C.POINTER(_i), C.POINTER(_i)]),
    "rg_compute_dos": (_i, [_vp, _d, _d, _d, _i, _pd, _pd]),
    "rg_kpm_density": (_i, [_pd, _i, _pd, _i64, _d, _i, _pd]),
    "rg_mfma_plan": (_i, [_i64, _pi64, _pi32, _pi32, _pi32, C.POINTER(C.c_int32)]),
    "rg_bsr_info": (_i, [_vp, _i, _pi64, C.POINTER(_i), _pi64, C.POINTER(_i)]),
    "rg_download_bsr": (_i, [_vp, _i, _pi64, _pi32, _pd]),
    "rg_sk_assemble": (_i, [_vp, _i, C.POINTER(SkTableC), _i64, _pi32, _pd, _pd, _i64, _pi32, _pi32, _pd,
                            _pi64]),
    "rg_sk_velocity": (_i, [_vp, C.POINTER(SkTableC), _i64, _pi32, _pd, _pd, _i64, _pi32, _pi32, _pd]),
    "rg_kubo_greenwood": (_i, [_vp, _i, _i, _d, _pd, _pd, _pd]),
    "rg_dc_spectrum": (_i, [_vp, _i, _i, _i, _d, _pd, _pd, _pd]),
    "rg_dc_spectrum_tangent": (_i, [_vp, _i, _i, _i, _d, _pd, _pd, _pd]),
    "rg_dc_spectrum_ex": (_i, [_vp, _i, _i, _i, _d, _i, _pd, _pd, _pd]),
    "rg_isqrt_patches": (_i, [_vp, _d, _i, C.POINTER(_i), _pd]),
    "rg_estimate_bounds": (_i, [_vp, _i, _i, _u64, _d, _pd, _pd]),
    "rg_set_checkpointing": (_i, [_vp, _i]),
    "rg_kpm_moments_backward": (_i, [_vp, _pd, _pd]),
    "rg_kpm_moments_backward_ex": (_i, [_vp, _pd, _pd, _pd, _pd]),
    "rg_cancel": (_i, [_vp]),
    "rg_set_profiling": (_i, [_vp, _i]),
    "rg_kernel_stats": (_i, [_vp, _i, _pd, _pi64, _pd, _pd]),
}

_lib = None


class RsatbGpuError(RuntimeError):
    """A failed call into the gfx950 library (maps rsatb's std::runtime_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def load() -> C.CDLL:
    """Loads the library once. Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -m rsatb_amd.build). "
            "rsatb_amd has no CPU fallback."
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_pd)


def _check(ctx_handle, rc: int, what: str) -> None:
    if rc == RG_OK:
        return
    msg = load().rg_last_error(ctx_handle).decode(errors="replace") if ctx_handle else ""
    if rc == RG_ERR_CANCELLED:
        raise KeyboardInterrupt(f"{what}: {msg or 'interrupted'}")
    raise RsatbGpuError(rc, f"{what} failed (status {rc}): {msg}")


def dos_plan(emin: float, emax: float, de: float) -> tuple[int, int]:
    npts, nmom = C.c_int(0), C.c_int(0)
    rc = load().rg_dos_plan(emin, emax, de, C.byref(npts), C.byref(nmom))
    if rc != RG_OK:
        raise RsatbGpuError(rc, "dos needs finite emin < emax and a positive energy resolution dE")
    return npts.value, nmom.value


def kpm_density(mu: np.ndarray, x: np.ndarray, edge_inset: float, kernel: int = KERNEL_JACKSON) -> np.ndarray:
    """evaluate_kpm_density (rsatb/calc/kpm.hpp:206-241), host FP64 in the library."""
    mu = np.ascontiguousarray(np.asarray(mu, dtype=np.complex128))
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    rho = np.zeros(x.shape[0], dtype=np.float64)
    mu_f = mu.view(np.float64)
    rc = load().rg_kpm_density(_dptr(mu_f), int(mu.shape[0]), _dptr(x), int(x.shape[0]), float(edge_inset), int(kernel), _dptr(rho))
    if rc != RG_OK:
        raise RsatbGpuError(rc, "kpm_density: bad arguments")
    return rho


def mfma_plan(row_ptr: np.ndarray, col_ind: np.ndarray) -> tuple[np.ndarray, np.ndarray, int]:
    """The MFMA SpMM plan of a square block pattern (rg_mfma_plan): block rows
    in processing order, the stored block behind every position, and the
    region size. Host only."""
    rp = np.ascontiguousarray(np.asarray(row_ptr, dtype=np.int64))
    ci = np.ascontiguousarray(np.asarray(col_ind, dtype=np.int32))
    n = int(rp.shape[0]) - 1
    sched = np.zeros(max(n, 0), dtype=np.int32)
    kperm = np.zeros(ci.shape[0], dtype=np.int32)
    sreg = C.c_int32(0)
    rc = load().rg_mfma_plan(n, rp.ctypes.data_as(_pi64), ci.ctypes.data_as(_pi32), sched.ctypes.data_as(_pi32),
                             kperm.ctypes.data_as(_pi32), C.byref(sreg))
    if rc != RG_OK:
        raise RsatbGpuError(rc, "mfma_plan: bad pattern")
    return sched, kperm, sreg.value


class Context:
    """One device context: operands, samples and the moment loop on a stream."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        rc = lib.rg_create(int(device), C.byref(h))
        if rc != RG_OK:
            raise RsatbGpuError(rc, f"rg_create(device={device}) failed: no usable HIP device")
        self._h = h
        self._lib = lib
        self.rows = None
        self.R = None

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.rg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def mem_info(self) -> tuple[int, int]:
        """(free, total) bytes of this context's device (rg_mem_info)."""
        free, total = C.c_size_t(), C.c_size_t()
        _check(self._h, self._lib.rg_mem_info(self._h, C.byref(free), C.byref(total)), "rg_mem_info")
        return int(free.value), int(total.value)

    # operands -------------------------------------------------------
    def upload_bsr(self, slot: int, bsr) -> None:
        vals = np.ascontiguousarray(bsr.values)
        if vals.dtype == np.float32:
            dtype = F32  # widened to FP64 by the library
        elif vals.dtype == np.complex64:
            dtype = C64
        elif np.iscomplexobj(vals):
            vals = vals.astype(np.complex128, copy=False)
            dtype = C128
        else:
            vals = vals.astype(np.float64, copy=False)
            dtype = F64
        rp = np.ascontiguousarray(bsr.row_ptr, dtype=np.int64)
        ci = np.ascontiguousarray(bsr.col_ind, dtype=np.int32)
        rc = self._lib.rg_upload_bsr(
            self._h, int(slot), int(bsr.n_block_rows), int(bsr.block_dim),
            rp.ctypes.data_as(_pi64), ci.ctypes.data_as(_pi32), vals.ctypes.data_as(C.c_void_p), dtype,
        )
        _check(self._h, rc, "rg_upload_bsr")

    def bsr_info(self, slot: int) -> dict:
        nbr, b, nnzb, cplx = C.c_int64(), C.c_int(), C.c_int64(), C.c_int()
        rc = self._lib.rg_bsr_info(self._h, int(slot), C.byref(nbr), C.byref(b), C.byref(nnzb), C.byref(cplx))
        _check(self._h, rc, "rg_bsr_info")
        return {"n_block_rows": nbr.value, "block_dim": b.value, "nnzb": nnzb.value, "is_complex": bool(cplx.value)}

    def download_bsr(self, slot: int):
        """A slot's operand in the upload layout (rsatb_amd.bsr.BSR)."""
        from .bsr import BSR

        info = self.bsr_info(slot)
        nbr, b, nnzb = info["n_block_rows"], info["block_dim"], info["nnzb"]
        rp = np.zeros(nbr + 1, dtype=np.int64)
        ci = np.zeros(max(nnzb, 1), dtype=np.int32)
        vals = np.zeros((max(nnzb, 1), b, b), dtype=np.complex128 if info["is_complex"] else np.float64)
        rc = self._lib.rg_download_bsr(self._h, int(slot), rp.ctypes.data_as(_pi64), ci.ctypes.data_as(_pi32),
                                       _dptr(vals.view(np.float64)))
        _check(self._h, rc, "rg_download_bsr")
        return BSR(nbr, b, rp, ci[:nnzb], vals[:nnzb])

    def sk_assemble(self, slot: int, table_arrays: dict, atom_type, pos, cell, edge_index, edge_shift,
                    kpoint=None) -> int:
        """Builds H(k) or S(k) of an SK table on the device into `slot`
        (table_arrays: rsatb_amd.sktb.SkTable.arrays(which); edge_index as
        flattened (src, dst) pairs). Returns the stored block count."""
        return self._sk_call(False, slot, table_arrays, atom_type, pos, cell, edge_index, edge_shift, kpoint)

    def sk_velocity(self, table_arrays: dict, atom_type, pos, cell, edge_index, edge_shift, kpoint=None) -> None:
        """Bond-moment velocity factors B_x, B_y, B_z (V = i B) of an
        orthogonal SK model into SLOT_VX .. SLOT_VZ, at Gamma (real) or at a
        fractional kpoint in the gauge of sk_assemble's H(k) (complex)."""
        self._sk_call(True, SLOT_VX, table_arrays, atom_type, pos, cell, edge_index, edge_shift, kpoint)

    def kubo_greenwood(self, n_moments: int, kernel: int = KERNEL_JACKSON, prefactor: float = 1.0,
                       with_moments: bool = False):
        """(sigma [M, 3] for xx, yy, zz on ascending nodes, energies [M]
        [, mu [3, M, M] complex])."""
        M = int(n_moments)
        sigma = np.zeros((M, 3))
        energy = np.zeros(M)
        mu = np.zeros((3, M, M), dtype=np.complex128) if with_moments else None
        rc = self._lib.rg_kubo_greenwood(self._h, M, int(kernel), float(prefactor), _dptr(sigma), _dptr(energy),
                                         _dptr(mu.view(np.float64)) if with_moments else None)
        _check(self._h, rc, "rg_kubo_greenwood")
        return (sigma, energy, mu) if with_moments else (sigma, energy)

    def dc_spectrum(self, n_moments: int, formulation: int = 0, kernel: int = KERNEL_JACKSON,
                    prefactor: float = 1.0, with_moments: bool = False, precision: int = PREC_FP64):
        """(sigma [M, 6] in (xx, xy, xz, yy, yz, zz), energies [M] [, mu
        [pairs, M, M]]); formulation 0 Greenwood, 1 Bastin; precision
        PREC_FP32 runs the recurrence in FP32 with FP64 moment accumulation."""
        M = int(n_moments)
        sigma = np.zeros((M, 6))
        energy = np.zeros(M)
        mu = np.zeros((6 if formulation == 1 else 3, M, M), dtype=np.complex128) if with_moments else None
        rc = self._lib.rg_dc_spectrum_ex(self._h, M, int(formulation), int(kernel), float(prefactor), int(precision),
                                         _dptr(sigma), _dptr(energy),
                                         _dptr(mu.view(np.float64)) if with_moments else None)
        _check(self._h, rc, "rg_dc_spectrum_ex")
        return (sigma, energy, mu) if with_moments else (sigma, energy)

    def dc_spectrum_tangent(self, n_moments: int, formulation: int = 0, kernel: int = KERNEL_JACKSON,
                            prefactor: float = 1.0):
        """(sigma [M, 6], dsigma [M, 6], energies [M]): the spectrum and its
        derivative along the tangent matrices in SLOT_DH, SLOT_DVX..DVZ."""
        M = int(n_moments)
        sigma = np.zeros((M, 6))
        dsigma = np.zeros((M, 6))
        energy = np.zeros(M)
        rc = self._lib.rg_dc_spectrum_tangent(self._h, M, int(formulation), int(kernel), float(prefactor),
                                              _dptr(sigma), _dptr(dsigma), _dptr(energy))
        _check(self._h, rc, "rg_dc_spectrum_tangent")
        return sigma, dsigma, energy

    def _sk_call(self, velocity: bool, slot: int, table_arrays: dict, atom_type, pos, cell, edge_index, edge_shift,
                 kpoint):
        a = table_arrays
        keep = {}

        def i32(name):
            keep[name] = np.ascontiguousarray(a[name], dtype=np.int32)
            return keep[name].ctypes.data_as(_pi32)

        def f64(name):
            keep[name] = np.ascontiguousarray(a[name], dtype=np.float64)
            return keep[name].ctypes.data_as(_pd)

        t = SkTableC(int(a["n_types"]), int(a["norb"]), int(a["n_channels"]), i32("orb_l"), i32("orb_m"),
                     f64("onsite"), f64("pair_cutoff"), i32("ch_row_type"), i32("ch_col_type"), i32("ch_row_l"),
                     i32("ch_col_l"), i32("ch_bond"), i32("ch_kind"), i32("ch_off"), i32("ch_n"),
                     f64("ch_r_cut"), f64("ch_lam"), f64("ch_r_on"), f64("rad_x"), f64("rad_y"), f64("rad_y2"),
                     int(np.asarray(a["rad_y"]).size))
        types = np.ascontiguousarray(atom_type, dtype=np.int32).reshape(-1)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        ei = np.ascontiguousarray(edge_index, dtype=np.int32).reshape(-1)
        sh = np.ascontiguousarray(edge_shift, dtype=np.int32).reshape(-1)
        if ei.size % 2 or sh.size != 3 * (ei.size // 2):
            raise ValueError("edge_index needs (src, dst) pairs and edge_shift three integers per edge")
        if pos.shape[0] != types.size:
            raise ValueError("pos and atom_type describe different atom counts")
        kp = None if kpoint is None else np.ascontiguousarray(kpoint, dtype=np.float64).reshape(3)
        if velocity:
            rc = self._lib.rg_sk_velocity(self._h, C.byref(t), int(types.size), types.ctypes.data_as(_pi32),
                                          _dptr(pos), _dptr(cell), int(ei.size // 2), ei.ctypes.data_as(_pi32),
                                          sh.ctypes.data_as(_pi32), None if kp is None else _dptr(kp))
            _check(self._h, rc, "rg_sk_velocity")
            return None
        nnzb = C.c_int64()
        rc = self._lib.rg_sk_assemble(self._h, int(slot), C.byref(t), int(types.size), types.ctypes.data_as(_pi32),
                                      _dptr(pos), _dptr(cell), int(ei.size // 2), ei.ctypes.data_as(_pi32),
                                      sh.ctypes.data_as(_pi32), None if kp is None else _dptr(kp),
                                      C.byref(nnzb))
        _check(self._h, rc, "rg_sk_assemble")
        return nnzb.value

    def isqrt_patches(self, tol: float = 1e-12, maxit: int = 60) -> tuple[int, float]:
        """Z, Z^H from the overlap in SLOT_S (graph_inverse_sqrt seed)."""
        it, res = C.c_int(), C.c_double()
        rc = self._lib.rg_isqrt_patches(self._h, float(tol), int(maxit), C.byref(it), C.byref(res))
        _check(self._h, rc, "rg_isqrt_patches")
        return it.value, res.value

    def estimate_bounds(self, form: int, steps: int = 200, seed: int = 20240917, padding: float = 0.5):
        lo, hi = C.c_double(), C.c_double()
        rc = self._lib.rg_estimate_bounds(self._h, int(form), int(steps), C.c_uint64(int(seed) & (2**64 - 1)),
                                          float(padding), C.byref(lo), C.byref(hi))
        _check(self._h, rc, "rg_estimate_bounds")
        return lo.value, hi.value

    def set_operator(self, form: int, center: float, half_width: float) -> None:
        _check(self._h, self._lib.rg_set_operator(self._h, int(form), float(center), float(half_width)), "rg_set_operator")

    def set_samples(self, n_rows: int, R: int, kind: int, seed: int = 20240917, data: np.ndarray | None = None) -> None:
        ptr = None
        if kind == SAMPLES_UPLOAD_REAL:
            data = np.ascontiguousarray(data, dtype=np.float64).reshape(n_rows, R)
            ptr = _dptr(data)
        elif kind == SAMPLES_UPLOAD_COMPLEX:
            data = np.ascontiguousarray(data, dtype=np.complex128).reshape(n_rows, R)
            ptr = _dptr(data.view(np.float64))
        rc = self._lib.rg_set_samples(self._h, int(n_rows), int(R), int(kind), C.c_uint64(int(seed) & (2**64 - 1)), ptr)
        _check(self._h, rc, "rg_set_samples")
        self.rows, self.R = int(n_rows), int(R)

    def get_samples(self, as_complex: bool = True) -> np.ndarray:
        if as_complex:
            out = np.zeros((self.rows, self.R), dtype=np.complex128)
            rc = self._lib.rg_get_samples(self._h, _dptr(out.view(np.float64)), 1)
        else:
            out = np.zeros((self.rows, self.R), dtype=np.float64)
            rc = self._lib.rg_get_samples(self._h, _dptr(out), 0)
        _check(self._h, rc, "rg_get_samples")
        return out

    # operator application --------------------------------------------
    def _vec_call(self, fn, *head, x: np.ndarray):
        x = np.asarray(x)
        cplx = np.iscomplexobj(x)
        if x.ndim == 1:
            x = x[:, None]
        R = x.shape[1]
        if cplx:
            xc = np.ascontiguousarray(x, dtype=np.complex128)
            y = np.zeros_like(xc)
            rc = fn(self._h, *head, R, _dptr(xc.view(np.float64)), _dptr(y.view(np.float64)), 1)
        else:
            xc = np.ascontiguousarray(x, dtype=np.float64)
            y = np.zeros_like(xc)
            rc = fn(self._h, *head, R, _dptr(xc), _dptr(y), 0)
        return rc, y

    def apply_op(self, x: np.ndarray) -> np.ndarray:
        """(A x - center x)/half_width (Chebyshev::apply_op, NORMAL)."""
        rc, y = self._vec_call(self._lib.rg_apply_op, x=x)
        _check(self._h, rc, "rg_apply_op")
        return y

    def bsr_mult(self, slot: int, x: np.ndarray) -> np.ndarray:
        rc, y = self._vec_call(self._lib.rg_bsr_mult, int(slot), x=x)
        _check(self._h, rc, "rg_bsr_mult")
        return y

    # the hot loop ------------------------------------------------------
    def kpm_moments(self, n_moments: int, precision: int = PREC_FP64, doubling: int = DOUBLING_AUTO,
                    per_column: bool = False):
        mu = np.zeros(n_moments, dtype=np.complex128)
        cols = np.zeros((n_moments, self.R), dtype=np.complex128) if per_column else None
        rc = self._lib.rg_kpm_moments(
            self._h, int(n_moments), int(precision), int(doubling), _dptr(mu.view(np.float64)),
            _dptr(cols.view(np.float64)) if per_column else None,
        )
        _check(self._h, rc, "rg_kpm_moments")
        return (mu, cols) if per_column else mu

    def kpm_moments_noalloc(self, n_moments: int, precision: int = PREC_FP64, doubling: int = DOUBLING_AUTO,
                            per_column: bool = False):
        """rg_kpm_moments_noalloc: the same moments with no device allocation."""
        mu = np.zeros(n_moments, dtype=np.complex128)
        cols = np.zeros((n_moments, self.R), dtype=np.complex128) if per_column else None
        rc = self._lib.rg_kpm_moments_noalloc(
            self._h, int(n_moments), int(precision), int(doubling), _dptr(mu.view(np.float64)),
            _dptr(cols.view(np.float64)) if per_column else None,
        )
        _check(self._h, rc, "rg_kpm_moments_noalloc")
        return (mu, cols) if per_column else mu

    def kpm_pdos_moments(self, n_moments: int, group_of_row, n_groups: int, precision: int = PREC_FP64,
                         doubling: int = DOUBLING_AUTO):
        """kpm_pdos_moments (rsatb/calc/pdos.hpp:149-266): (total [M], group [n_groups, M])."""
        grp = np.ascontiguousarray(np.asarray(group_of_row, dtype=np.int32).reshape(-1))
        total = np.zeros(n_moments, dtype=np.complex128)
        group = np.zeros((max(int(n_groups), 0), n_moments), dtype=np.complex128)
        rc = self._lib.rg_kpm_pdos_moments(
            self._h, int(n_moments), int(precision), int(doubling), int(grp.size),
            grp.ctypes.data_as(_pi32), int(n_groups), _dptr(total.view(np.float64)),
            _dptr(group.view(np.float64)) if group.size else None,
        )
        _check(self._h, rc, "rg_kpm_pdos_moments")
        return total, group

    def set_checkpointing(self, stride: int) -> None:
        _check(self._h, self._lib.rg_set_checkpointing(self._h, int(stride)), "rg_set_checkpointing")

    def kpm_moments_backward(self, dL_dmu: np.ndarray) -> np.ndarray:
        """dL/dH per stored entry of H ([nnzb, b, b], upload layout)."""
        info = self.bsr_info(SLOT_H)
        g = np.ascontiguousarray(np.asarray(dL_dmu, dtype=np.float64))
        out = np.zeros((info["nnzb"], info["block_dim"], info["block_dim"]))
        rc = self._lib.rg_kpm_moments_backward(self._h, _dptr(g), _dptr(out))
        _check(self._h, rc, "rg_kpm_moments_backward")
        return out

    def kpm_moments_backward_ex(self, dL_dmu: np.ndarray, want_h: bool = True, want_z: bool = False,
                                want_s: bool = False) -> dict:
        """Per-slot gradients {"H", "Z", "S"} ([nnzb, b, b] each, upload layout)
        of the requested slots (rg_kpm_moments_backward_ex)."""
        g = np.ascontiguousarray(np.asarray(dL_dmu, dtype=np.float64))
        outs = {}
        ptrs = []
        for name, slot, want in (("H", SLOT_H, want_h), ("Z", SLOT_Z, want_z), ("S", SLOT_S, want_s)):
            if want:
                info = self.bsr_info(slot)
                outs[name] = np.zeros((info["nnzb"], info["block_dim"], info["block_dim"]))
                ptrs.append(_dptr(outs[name]))
            else:
                ptrs.append(None)
        rc = self._lib.rg_kpm_moments_backward_ex(self._h, _dptr(g), *ptrs)
        _check(self._h, rc, "rg_kpm_moments_backward_ex")
        return outs

    def compute_dos(self, emin: float, emax: float, de: float, precision: int = PREC_FP64):
        npts, _ = dos_plan(emin, emax, de)
        dos = np.zeros(npts)
        omega = np.zeros(npts)
        rc = self._lib.rg_compute_dos(self._h, float(emin), float(emax), float(de), int(precision), _dptr(dos), _dptr(omega))
        _check(self._h, rc, "rg_compute_dos")
        return dos, omega

    def cancel(self) -> None:
        _check(self._h, self._lib.rg_cancel(self._h), "rg_cancel")

    # measurement --------------------------------------------------------
    def set_profiling(self, on: bool) -> None:
        _check(self._h, self._lib.rg_set_profiling(self._h, 1 if on else 0), "rg_set_profiling")

    def kernel_stats(self, kclass: int) -> dict:
        avg, cnt, by, fl = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
        rc = self._lib.rg_kernel_stats(self._h, int(kclass), C.byref(avg), C.byref(cnt), C.byref(by), C.byref(fl))
        _check(self._h, rc, "rg_kernel_stats")
        return {"avg_ms": avg.value, "count": cnt.value, "bytes": by.value, "flops": fl.value}
