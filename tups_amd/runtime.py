"""ctypes binding of ``libtups_gpu.so`` (the C ABI declared in include/tups_gpu.h).

The class :class:`TupsContext` mirrors the reference's hot-path routines by
name (``BSE``, ``calc_PhiR``, ``calc_M_*``, ``calculate_PAE``) and keeps their
argument meaning: the routines take no arrays — they act on the vertex arrays
held by the context, exactly as the Fortran routines act on the module
variables of ``parquet_ini`` (SRC/parquet_ini.f90:128-137).  Arrays cross the
boundary in the reference's Fortran layouts (numpy ``order='F'``).

There is no CPU fallback: if the HIP library is missing or no GPU is visible
the calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtups_gpu.so"

GAMMA, F, CHI, PHIR, M, LAMBDA = range(6)
(K_BSE_LADDER, K_PHIR, K_BUILD_M, K_PAE, K_GAMMA_UPDATE, K_PAE_GATHER, K_SPLIT_PREP, K_SPLIT_GEMM,
 K_BSE_CLOSED, K_CALC_CHI, K_BSE_BATCH, K_SDE, K_GTAU, K_CONV, K_SUS, K_BSE_IR) = range(16)
CH_D, CH_M, CH_S, CH_T = range(4)
BSE_LADDER_FP64 = 0
BSE_LADDER_SPLIT = 1
BSE_CLOSED_LU = 2
BSE_CLOSED_IR = 3
M_NAMES = ("phbar_to_ph", "pp_to_ph", "ph_to_pp", "phbar_to_pp")

ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ESTATE", -5: "ENONFINITE", -6: "ENOCONV"}


class TupsError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({ERRORS.get(code, code)}): {msg}")
        self.code = code


class Dims(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nf", C.c_int32), ("nl", C.c_int32), ("nr", C.c_int32),
                ("nq", C.c_int32), ("reserved", C.c_int32), ("beta", C.c_double),
                ("f_damping", C.c_double)]


class Prms(C.Structure):
    _fields_ = ([("beta", C.c_double), ("u", C.c_double), ("n_particle", C.c_double)]
                + [(n, C.c_int32) for n in ("nx", "nf", "neighbours", "readinpick", "ngrain", "f_range", "ite_max",
                                            "dga", "use_sig_dmft", "use_u2", "save_f", "oneshot", "old", "sde_old")]
                + [("f_damping", C.c_double), ("pick_file", C.c_char * 104), ("fname_l", C.c_char * 104),
                   ("fname_sig_dmft", C.c_char * 104)])


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nx", "ny", "nf", "nl", "nz", "nr", "nw", "nibz", "nq",
                                         "nc", "full_ibz", "reserved")] + [("alpha", C.c_double)]


class ConvInfo(C.Structure):
    _fields_ = [("sigma_avg", C.c_double), ("sigma_max", C.c_double), ("kx", C.c_int32), ("ky", C.c_int32),
                ("nu", C.c_int32), ("converged", C.c_int32), ("f_max", C.c_double * 4),
                ("f_rel_err", C.c_double * 4)]


_P = C.POINTER
_D = _P(C.c_double)
_I = _P(C.c_int32)
_V = C.c_void_p

SIGNATURES = {
    "tups_last_error": ([], C.c_char_p),
    "tups_build_tables": ([_P(Dims), _D, _D, _I, _I, _D, _D], C.c_int),
    "tups_create": ([_P(Dims), _P(_V)], C.c_int),
    "tups_destroy": ([_V], C.c_int),
    "tups_query": ([_V, _P(Sizes)], C.c_int),
    "tups_get_tables": ([_V, _D, _D, _I, _I, _D, _D], C.c_int),
    "tups_set_tables": ([_V, _D, _D, _I, _I, _D, _D], C.c_int),
    "tups_get_q_map": ([_V, _I, _I, _I], C.c_int),
    "tups_set_q_map": ([_V, _I, _I, _I], C.c_int),
    "tups_upload": ([_V, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_download": ([_V, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_upload_chi_dense": ([_V, C.c_int32, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_bse": ([_V, C.c_int32], C.c_int),
    "tups_set_refinement": ([_V, C.c_double, C.c_int32], C.c_int),
    "tups_bse_solve_info": ([_V, C.POINTER(C.c_int32), _D], C.c_int),
    "tups_set_split": ([_V, C.c_int32, C.c_int32], C.c_int),
    "tups_calc_chi": ([_V, _D, _D, C.c_double, C.c_int32], C.c_int),
    "tups_sde": ([_V, _D, _D, C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_int32, _D], C.c_int),
    "tups_set_sigma": ([_V, _D, _D, C.c_double, C.c_int32], C.c_int),
    "tups_get_sigma": ([_V, _D], C.c_int),
    "tups_iteration_full": ([_V, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double], C.c_int),
    "tups_gkw_grt": ([_V, C.c_double, C.c_double, C.c_int32, _D, C.POINTER(C.c_int32), _D], C.c_int),
    "tups_get_mu": ([_V, _D, C.POINTER(C.c_int32)], C.c_int),
    "tups_chi0": ([_V, _D, _D, _D, _D], C.c_int),
    "tups_conv": ([_V, _D, _D, C.c_double, C.c_double, _P(ConvInfo)], C.c_int),
    "tups_conv_noalloc": ([_V, _D, _D, C.c_double, C.c_double, _P(ConvInfo)], C.c_int),
    "tups_read_prms": ([C.c_char_p, _P(Prms)], C.c_int),
    "tups_calc_sus": ([_V, C.c_double, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_prms_dims": ([_P(Prms), _P(Dims)], C.c_int),
    "tups_iteration": ([_V, C.c_int32, C.c_int32], C.c_int),
    "tups_batch_create": ([C.c_int32, C.c_int32, _I, _I, C.c_double, C.c_double, _P(_V)], C.c_int),
    "tups_batch_destroy": ([_V], C.c_int),
    "tups_batch_query": ([_V, _P(C.c_int64), _P(C.c_int64), _I], C.c_int),
    "tups_batch_upload": ([_V, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_batch_download": ([_V, C.c_int32, C.c_int32, _D], C.c_int),
    "tups_batch_bse": ([_V, _P(C.c_float)], C.c_int),
    "tups_batch_bse_split": ([_V, C.c_int32, _P(C.c_float)], C.c_int),
    "tups_batch_sync": ([_V], C.c_int),
    "tups_batch_autotune": ([_V, _I, _P(C.c_float)], C.c_int),
    "tups_batch_set_variant": ([_V, C.c_int32], C.c_int),
    "tups_batch_set_product_forms": ([_V, C.c_int32], C.c_int),
    "tups_set_product_forms": ([_V, C.c_int32], C.c_int),
    "tups_uneven_create": ([C.c_int32, C.c_int32, _I, _I, C.c_double, C.c_double, _P(_V)], C.c_int),
    "tups_uneven_destroy": ([_V], C.c_int),
    "tups_uneven_parts": ([_V, _P(_V), _P(_V)], C.c_int),
    "tups_uneven_iteration": ([_V, C.c_int32], C.c_int),
    "tups_uneven_sync": ([_V], C.c_int),
    "tups_phir": ([_V], C.c_int),
    "tups_build_m": ([_V], C.c_int),
    "tups_pae": ([_V], C.c_int),
    "tups_gamma_update": ([_V], C.c_int),
    "tups_autotune": ([_V, C.c_int32, _I, _P(C.c_float)], C.c_int),
    "tups_variant_name": ([C.c_int32, C.c_int32], C.c_char_p),
    "tups_set_variant": ([_V, C.c_int32, C.c_int32], C.c_int),
    "tups_variant_count": ([C.c_int32], C.c_int),
    "tups_sync": ([_V], C.c_int),
    "tups_event_record": ([_V, C.c_int32], C.c_int),
    "tups_event_elapsed": ([_V, C.c_int32, C.c_int32, _P(C.c_float)], C.c_int),
    "tups_set_timing": ([_V, C.c_int32], C.c_int),
    "tups_get_timing": ([_V, C.c_int32, _D, _P(C.c_int64)], C.c_int),
    "tups_check_finite": ([_V, C.c_int32, C.c_int32], C.c_int),
    "tups_synth_vertex": ([_V, C.c_int32, C.c_int32, C.c_uint64, C.c_double, C.c_double], C.c_int),
    "tups_copy": ([_V, C.c_int32, C.c_int32, C.c_int32], C.c_int),
}

_lib = None


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Load the HIP library once; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m tups_amd.build` "
                           "(there is no CPU fallback for the hot path)")
    lib = C.CDLL(str(p))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


H5_LIB_PATH = Path(__file__).resolve().parent / "libtups_h5.so"
_h5 = None


def load_h5_library() -> C.CDLL:
    """libtups_h5.so (the HDF5 vertex dump), loaded after libtups_gpu.so."""
    global _h5
    if _h5 is None:
        load_library()
        if not H5_LIB_PATH.exists():
            raise RuntimeError(f"{H5_LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(needs the HDF5 C library)")
        lib = C.CDLL(str(H5_LIB_PATH))
        lib.tups_h5_write_vertex.argtypes = [C.c_char_p, C.c_char_p, _D, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        lib.tups_h5_write_vertex.restype = C.c_int
        lib.tups_h5_write_f.argtypes = [_V, C.c_char_p]
        lib.tups_h5_write_f.restype = C.c_int
        lib.tups_h5_last_error.argtypes = []
        lib.tups_h5_last_error.restype = C.c_char_p
        _h5 = lib
    return _h5


def h5_write_vertex(path, dset: str, data, create: bool = True) -> None:
    """hdf5_write_data of one (Nz, Nz, Nb) vertex as reshape(., (Nf, Nl, Nf, Nl, Nb)); data is
    passed with its (nf, nl, nb) as a (nf, nl, nf, nl, nb) Fortran-order array."""
    a = np.asfortranarray(np.asarray(data, dtype=np.complex128))
    assert a.ndim == 5 and a.shape[0] == a.shape[2] and a.shape[1] == a.shape[3]
    lib = load_h5_library()
    rc = lib.tups_h5_write_vertex(str(path).encode(), dset.encode(), a.ctypes.data_as(_D), a.shape[0], a.shape[1],
                                  a.shape[4], 1 if create else 0)
    if rc != 0:
        raise TupsError(rc, "tups_h5_write_vertex", lib.tups_h5_last_error().decode())


_PRMS_BOOL = ("readinpick", "dga", "use_sig_dmft", "use_u2", "save_f", "oneshot", "old", "sde_old")


def read_prms(path) -> dict:
    """The reference's run-parameter file (set_parameters, SRC/parquet_ini.f90:289-329) as a
    dict of its 21 values, plus the context sizes it implies under "dims" (nx, nf, nl, beta,
    f_damping; an unsupported neighbours value leaves "dims" as None with the reason)."""
    lib = load_library()
    p = Prms()
    _check(lib.tups_read_prms(str(path).encode(), C.byref(p)), "tups_read_prms")
    out = {}
    for name, _ in Prms._fields_:
        v = getattr(p, name)
        if isinstance(v, bytes):
            v = v.decode()
        elif name in _PRMS_BOOL:
            v = bool(v)
        out[name] = v
    d = Dims()
    rc = lib.tups_prms_dims(C.byref(p), C.byref(d))
    if rc == 0:
        out["dims"] = {"nx": d.nx, "nf": d.nf, "nl": d.nl, "beta": d.beta, "f_damping": d.f_damping}
    else:
        out["dims"] = None
        out["dims_error"] = lib.tups_last_error().decode()
    return out


def _check(rc: int, where: str):
    if rc != 0:
        raise TupsError(rc, where, load_library().tups_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_I)


def make_dims(nx, nf, nl, nr=0, nq=0, beta=2.0, f_damping=0.5) -> Dims:
    return Dims(nx, nf, nl, nr, nq, 0, beta, f_damping)


def build_tables(nx, nf, nl, nr=0):
    """Host-only table build (no GPU needed): FF, FF_inv, signs, reps, expSqR, expqSR."""
    lib = load_library()
    d = make_dims(nx, nf, nl, nr)
    nr_ = nr if nr > 0 else {1: 1, 5: 4, 9: 6}.get(nl, 1)
    nxi = (nx + 2) // 2
    nibz = nxi * (nxi + 1) // 2
    ff = np.zeros((nx, nx, nl), order="F")
    ffi = np.zeros((8, nx, nx, nl), order="F")
    signs = np.zeros(nl, dtype=np.int32)
    reps = np.zeros(nl, dtype=np.int32)
    esq = np.zeros((nibz, nr_, 8), dtype=np.complex128, order="F")
    eqs = np.zeros((nx * nx, nr_, 8), dtype=np.complex128, order="F")
    _check(lib.tups_build_tables(C.byref(d), _dptr(ff), _dptr(ffi), _iptr(signs), _iptr(reps),
                                 esq.ctypes.data_as(_D), eqs.ctypes.data_as(_D)), "tups_build_tables")
    return ff, ffi, signs, reps, esq, eqs


class TupsContext:
    """HBM-resident state of one rank's share of the TUPS vertex hot path."""

    def __init__(self, nx: int, nf: int, nl: int, nr: int = 0, nq: int = 0, beta: float = 2.0,
                 f_damping: float = 0.5, _handle=None):
        self.lib = load_library()
        self._owned = _handle is None
        if _handle is None:
            self._h = _V()
            d = make_dims(nx, nf, nl, nr, nq, beta, f_damping)
            _check(self.lib.tups_create(C.byref(d), C.byref(self._h)), "tups_create")
        else:
            self._h = _handle
        s = Sizes()
        _check(self.lib.tups_query(self._h, C.byref(s)), "tups_query")
        self.sizes = s
        self.nx, self.nf, self.nl, self.nz = s.nx, s.nf, s.nl, s.nz
        self._ngrain = 1
        self.nr, self.nw, self.nibz, self.nq, self.nc = s.nr, s.nw, s.nibz, s.nq, s.nc
        self.alpha = s.alpha
        self.beta = beta
        self.f_damping = f_damping

    # --- lifetime -------------------------------------------------------------------
    def close(self):
        if self._h:
            if self._owned:
                self.lib.tups_destroy(self._h)
            self._h = _V()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- shapes ---------------------------------------------------------------------
    def shape(self, which: int):
        nz, nl, nf = self.nz, self.nl, self.nf
        return {GAMMA: (nz, nz, self.nq), F: (nz, nz, self.nq), CHI: (nl, nl, nf, self.nq),
                PHIR: (nz * nz, self.nw, self.nr), M: (nl, nl, nl, nl, self.nibz, self.nr),
                LAMBDA: (nf, nf, self.nw)}[which]

    # --- transfers ------------------------------------------------------------------
    def upload(self, which: int, channel: int, arr: np.ndarray):
        a = np.asfortranarray(arr, dtype=np.complex128)
        if a.shape != self.shape(which):
            raise ValueError(f"shape {a.shape} != {self.shape(which)}")
        _check(self.lib.tups_upload(self._h, which, channel, a.ctypes.data_as(_D)), "tups_upload")

    def download(self, which: int, channel: int, out: np.ndarray = None) -> np.ndarray:
        """Device array -> host; into `out` (complex128, Fortran order, the array's
        shape) when given, as the drop-in fills the host mirrors in place."""
        if out is None:
            out = np.zeros(self.shape(which), dtype=np.complex128, order="F")
        elif out.dtype != np.complex128 or out.shape != self.shape(which) or not out.flags.f_contiguous:
            raise ValueError("out must be a Fortran-ordered complex128 array of the array's shape")
        _check(self.lib.tups_download(self._h, which, channel, out.ctypes.data_as(_D)), "tups_download")
        return out

    def upload_chi_dense(self, pp: int, dense: np.ndarray, q0: int = 0):
        a = np.asfortranarray(dense, dtype=np.complex128)
        if a.shape[:2] != (self.nz, self.nz):
            raise ValueError("dense bubble must be (Nz, Nz, nq)")
        _check(self.lib.tups_upload_chi_dense(self._h, pp, q0, a.shape[2], a.ctypes.data_as(_D)),
               "tups_upload_chi_dense")

    def q_map(self):
        qx = np.zeros(self.nq, dtype=np.int32)
        qy = np.zeros(self.nq, dtype=np.int32)
        qw = np.zeros(self.nq, dtype=np.int32)
        _check(self.lib.tups_get_q_map(self._h, _iptr(qx), _iptr(qy), _iptr(qw)), "tups_get_q_map")
        return qx, qy, qw

    def set_q_map(self, qx, qy, qw):
        qx, qy, qw = (np.ascontiguousarray(v, dtype=np.int32) for v in (qx, qy, qw))
        _check(self.lib.tups_set_q_map(self._h, _iptr(qx), _iptr(qy), _iptr(qw)), "tups_set_q_map")

    # --- the reference's routines -------------------------------------------------------
    def BSE(self, ite: int = 0, mode: int = BSE_LADDER_FP64):
        """BSE(ite), SRC/parquet_BSE.f90:19-135: G_r_LL <- Phi_r for r = d, m, s, t."""
        _check(self.lib.tups_bse(self._h, mode), "tups_bse")

    def set_refinement(self, tol: float = 1e-12, max_steps: int = 10):
        """Residual bound and step limit of the refined closed form (BSE_CLOSED_IR)."""
        _check(self.lib.tups_set_refinement(self._h, float(tol), int(max_steps)), "tups_set_refinement")

    def solve_info(self):
        """(refinement steps, largest |R|_F / |Gamma|_F) of the last closed-form BSE."""
        steps = C.c_int32(0)
        res = C.c_double(0.0)
        _check(self.lib.tups_bse_solve_info(self._h, C.byref(steps), C.byref(res)), "tups_bse_solve_info")
        return steps.value, res.value

    def set_split(self, moduli: int = 12, qbatch: int = 0):
        """CRT moduli and q points per pass of the split (INT8 MFMA) ladder."""
        _check(self.lib.tups_set_split(self._h, moduli, qbatch), "tups_set_split")

    def calc_chi(self, sigma, mu: float, ngrain: int = 1, sigma_h=None):
        """calc_chi for every q slot (SRC/parquet_chi.f90:438-523) into the resident bubbles.

        sigma: complex (Nx*Ny*Nf,) in the reference's order ((kx-1)*Ny + ky-1)*Nf + nu;
        sigma_h: complex (Nx, Ny) Hartree values used outside the box, or None for
        Sigma(k, Nf) as the reference sets them."""
        s = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128).ravel())
        assert s.size == self.nx * self.nx * self.nf
        sh = None
        if sigma_h is not None:
            sh = np.asfortranarray(np.asarray(sigma_h, dtype=np.complex128).reshape(self.nx, self.nx))
        _check(self.lib.tups_calc_chi(self._h, _dptr(s), _dptr(sh) if sh is not None else None, float(mu),
                                      int(ngrain)), "tups_calc_chi")
        self._ngrain = int(ngrain)

    def sde(self, sigma, mu: float, U: float, ngrain: int = 1, f_range: int = 0, use_u2: bool = False,
            sigma_h=None) -> np.ndarray:
        """calc_SDE (SRC/parquet_SDE.f90:27-258): Sigma (Nx*Ny*Nf,) over the whole zone
        from the resident G_r_LL and Lambda.  sigma / sigma_h / mu / ngrain build G as in
        calc_chi; use_u2 adds calc_order_2nd from the Grt of the last gkw_grt."""
        s = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128).ravel())
        assert s.size == self.nx * self.nx * self.nf
        sh = None
        if sigma_h is not None:
            sh = np.asfortranarray(np.asarray(sigma_h, dtype=np.complex128).reshape(self.nx, self.nx))
        out = np.zeros(self.nx * self.nx * self.nf, dtype=np.complex128)
        _check(self.lib.tups_sde(self._h, _dptr(s), _dptr(sh) if sh is not None else None, float(mu), int(ngrain),
                                 int(f_range), float(U), 1 if use_u2 else 0, _dptr(out)), "tups_sde")
        self._ngrain = int(ngrain)
        return out

    def set_sigma(self, sigma, mu: float, ngrain: int = 1, sigma_h=None):
        """Sigma, Sigma_H (None: Sigma(k, Nf)), mu and ngrain on the device for iteration_full."""
        s = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128).ravel())
        assert s.size == self.nx * self.nx * self.nf
        sh = None
        if sigma_h is not None:
            sh = np.asfortranarray(np.asarray(sigma_h, dtype=np.complex128).reshape(self.nx, self.nx))
        _check(self.lib.tups_set_sigma(self._h, _dptr(s), _dptr(sh) if sh is not None else None, float(mu),
                                       int(ngrain)), "tups_set_sigma")
        self._ngrain = int(ngrain)

    def get_sigma(self) -> np.ndarray:
        out = np.zeros(self.nx * self.nx * self.nf, dtype=np.complex128)
        _check(self.lib.tups_get_sigma(self._h, _dptr(out)), "tups_get_sigma")
        return out

    def iteration_full(self, U: float, mode: int = BSE_LADDER_FP64, graph: bool = True, f_range: int = 0,
                       use_u2: bool = False, n_particle: float = -1.0):
        """[calc_Gkw_Grt mu search when n_particle >= 0 ->] calc_chi -> BSE -> calc_PhiR ->
        calculate_PAE -> SDE [+ second order] -> Sigma update -> G = F - G on the device
        (SRC/loop.f90:67-222)."""
        _check(self.lib.tups_iteration_full(self._h, mode, 1 if graph else 0, int(f_range), float(U),
                                            1 if use_u2 else 0, float(n_particle)), "tups_iteration_full")

    def gkw_grt(self, U: float, n_particle: float, mu: float, cal_chi: bool = False, want_grt: bool = True):
        """calc_Gkw_Grt (SRC/parquet_chi.f90:18-210) from the device Sigma: returns
        (mu, steps, Grt) with Grt of shape (Nx Ng, Ny Ng, Nf) (None unless want_grt)."""
        m = C.c_double(float(mu))
        steps = C.c_int32(0)
        grt = None
        if want_grt:
            n1 = self.nx * self._ngrain
            grt = np.zeros((n1, n1, self.nf), dtype=np.complex128, order="F")
        _check(self.lib.tups_gkw_grt(self._h, float(U), float(n_particle), 1 if cal_chi else 0, C.byref(m),
                                     C.byref(steps), _dptr(grt) if grt is not None else None), "tups_gkw_grt")
        return m.value, steps.value, grt

    def conv(self, f_sigma: float = 1e-8, f_vert: float = 1e-8, sigma_old=None, sigma=None,
             noalloc: bool = False) -> dict:
        """conv() (SRC/parquet_check.f90:39-150) on the device: the Sigma pair from the host, or
        (both None) the pair the last iteration_full left on the device; F from the resident
        vertices; the previous F maxima are the context's.  noalloc calls tups_conv_noalloc
        (the library's fixed scratch array instead of a workspace allocated at the first call)."""
        so = sn = None
        if sigma is not None:
            so = np.ascontiguousarray(np.asarray(sigma_old, dtype=np.complex128).ravel())
            sn = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128).ravel())
        info = ConvInfo()
        name = "tups_conv_noalloc" if noalloc else "tups_conv"
        _check(getattr(self.lib, name)(self._h, _dptr(so) if so is not None else None,
                                       _dptr(sn) if sn is not None else None, float(f_sigma), float(f_vert),
                                       C.byref(info)), name)
        return {"sigma_avg": info.sigma_avg, "sigma_max": info.sigma_max, "kx": info.kx, "ky": info.ky,
                "nu": info.nu, "converged": bool(info.converged), "f_max": list(info.f_max),
                "f_rel_err": list(info.f_rel_err)}

    def calc_sus(self, U: float, f_range: int = 0, after_update: bool = False) -> np.ndarray:
        """calc_sus (SRC/parquet_sus.f90:18-362): (4 channels d, m, s, t; 4 form factors; Nq)."""
        out = np.zeros((4, 4, self.nq), dtype=np.complex128)
        _check(self.lib.tups_calc_sus(self._h, float(U), int(f_range), 1 if after_update else 0, _dptr(out)),
               "tups_calc_sus")
        return out

    def write_f(self, path) -> None:
        """write_F (SRC/parquet_plot.f90:237-276): the four resident F_r into F/F_d .. F/F_t of a
        new HDF5 file (libtups_h5.so)."""
        h5 = load_h5_library()
        rc = h5.tups_h5_write_f(self._h, str(path).encode())
        if rc != 0:
            raise TupsError(rc, "tups_h5_write_f", h5.tups_h5_last_error().decode())

    def chi0(self):
        """The bare bubbles of calc_Gkw_Grt with cal_chi from the device Grt:
        (Chi0_ph, Chi0_pp) of shape (Nred,) and (Chi0_ph_FF, Chi0_pp_FF) of shape (Nred, 4)."""
        nred = self.nibz * (self.nf // 2)
        ph = np.zeros(nred, dtype=np.complex128)
        pp = np.zeros(nred, dtype=np.complex128)
        phf = np.zeros((nred, 4), dtype=np.complex128, order="F")
        ppf = np.zeros((nred, 4), dtype=np.complex128, order="F")
        _check(self.lib.tups_chi0(self._h, _dptr(ph), _dptr(pp), _dptr(phf), _dptr(ppf)), "tups_chi0")
        return ph, pp, phf, ppf

    def get_mu(self):
        """(mu, steps) of the device search state."""
        m = C.c_double(0.0)
        steps = C.c_int32(0)
        _check(self.lib.tups_get_mu(self._h, C.byref(m), C.byref(steps)), "tups_get_mu")
        return m.value, steps.value

    def iteration(self, mode: int = BSE_LADDER_FP64, graph: bool = True):
        """BSE -> calc_PhiR -> calculate_PAE -> G = F - G (SRC/loop.f90:78-139,219-222);
        with graph=True the sequence is captured once and replayed as a HIP graph."""
        _check(self.lib.tups_iteration(self._h, mode, 1 if graph else 0), "tups_iteration")

    def calc_PhiR(self, ite: int = 0):
        """calc_PhiR(ite), SRC/parquet_PhiR.f90:35-109."""
        _check(self.lib.tups_phir(self._h), "tups_phir")

    def calc_M(self):
        """calc_M_phbar_to_ph / pp_to_ph / ph_to_pp / phbar_to_pp, SRC/parquet_EQMatrix.f90:21-360."""
        _check(self.lib.tups_build_m(self._h), "tups_build_m")

    calc_M_phbar_to_ph = calc_M_pp_to_ph = calc_M_ph_to_pp = calc_M_phbar_to_pp = calc_M

    def calculate_PAE(self):
        """calculate_PAE, SRC/parquet_EQOuter.f90:16-118."""
        _check(self.lib.tups_pae(self._h), "tups_pae")

    def gamma_update(self):
        """G_r_LL = F_r_LL - G_r_LL, SRC/loop.f90:219-222."""
        _check(self.lib.tups_gamma_update(self._h), "tups_gamma_update")

    def autotune(self, kernel: int = K_BSE_LADDER):
        """Pick the fastest tile variant on the resident inputs; returns (name, ms per launch)."""
        v = C.c_int32()
        ms = C.c_float()
        _check(self.lib.tups_autotune(self._h, kernel, C.byref(v), C.byref(ms)), "tups_autotune")
        return self.lib.tups_variant_name(kernel, v.value).decode(), float(ms.value)

    def variants(self, kernel: int = K_BSE_LADDER):
        """Variant ids (and names) of `kernel` usable with this context's Nl."""
        out = []
        for v in range(self.lib.tups_variant_count(kernel)):
            name = self.lib.tups_variant_name(kernel, v).decode()
            if kernel in (K_SPLIT_GEMM, K_BSE_CLOSED, K_BSE_IR):  # every variant serves every Nl
                out.append((v, name))
                continue
            if not name.startswith(f"nl{self.nl}_"):
                continue
            if kernel == K_PAE and "_nr" not in name and self.nr != 1:
                continue
            out.append((v, name))
        return out

    def set_variant(self, variant: int, kernel: int = K_BSE_LADDER):
        _check(self.lib.tups_set_variant(self._h, kernel, variant), "tups_set_variant")

    def set_product_forms(self, allow_three: bool):
        """Let the ladder autotuner pick the three-multiplication complex products."""
        _check(self.lib.tups_set_product_forms(self._h, 1 if allow_three else 0), "tups_set_product_forms")

    # --- stream / timing ----------------------------------------------------------------
    def sync(self):
        _check(self.lib.tups_sync(self._h), "tups_sync")

    def record(self, slot: int):
        _check(self.lib.tups_event_record(self._h, slot), "tups_event_record")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        _check(self.lib.tups_event_elapsed(self._h, a, b, C.byref(ms)), "tups_event_elapsed")
        return float(ms.value)

    def set_timing(self, on: bool):
        """Per-kernel HIP-event timing on the context stream (serialises launches)."""
        _check(self.lib.tups_set_timing(self._h, 1 if on else 0), "tups_set_timing")

    def kernel_time(self, kernel: int):
        """(total milliseconds, launches) accumulated for one kernel id."""
        ms = C.c_double()
        n = C.c_int64()
        _check(self.lib.tups_get_timing(self._h, kernel, C.byref(ms), C.byref(n)), "tups_get_timing")
        return float(ms.value), int(n.value)

    def check_finite(self, which: int, channel: int):
        _check(self.lib.tups_check_finite(self._h, which, channel), "tups_check_finite")

    def synth_vertex(self, which: int, channel: int, seed: int, diag: float, sigma: float):
        _check(self.lib.tups_synth_vertex(self._h, which, channel, seed, diag, sigma), "tups_synth_vertex")

    def copy(self, which_dst: int, which_src: int, channel: int):
        _check(self.lib.tups_copy(self._h, which_dst, which_src, channel), "tups_copy")


class TupsBatch:
    """Batching form: q points with their own (Nf_q, Nl_q); no reference counterpart.

    Vertex arrays are lists of per-q (D_q, D_q) complex matrices, bubbles lists
    of per-q (Nl_q, Nl_q, Nf_q) blocks; they are packed q after q for the C ABI.
    """

    def __init__(self, nx: int, nf, nl, beta: float = 2.0, f_damping: float = 0.5, _handle=None):
        self.lib = load_library()
        self.nf = [int(x) for x in nf]
        self.nl = [int(x) for x in nl]
        self.nq = len(self.nf)
        self.nx = nx
        self.alpha = 1.0 / (beta * nx * nx)
        self._owned = _handle is None
        if _handle is None:
            h = _V()
            nfa = np.ascontiguousarray(self.nf, dtype=np.int32)
            nla = np.ascontiguousarray(self.nl, dtype=np.int32)
            _check(self.lib.tups_batch_create(nx, self.nq, _iptr(nfa), _iptr(nla), beta, f_damping, C.byref(h)),
                   "tups_batch_create")
        else:
            h = _handle
        self._h = h
        ve, ce, it = C.c_int64(), C.c_int64(), C.c_int32()
        _check(self.lib.tups_batch_query(self._h, C.byref(ve), C.byref(ce), C.byref(it)), "tups_batch_query")
        self.vertex_elems, self.chi_elems, self.items = int(ve.value), int(ce.value), int(it.value)

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                self.lib.tups_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self):
        return [l * f for l, f in zip(self.nl, self.nf)]

    def _pack(self, blocks, n):
        out = np.empty(n, dtype=np.complex128)
        o = 0
        for b in blocks:
            flat = np.asarray(b, dtype=np.complex128).ravel(order="F")
            out[o:o + flat.size] = flat
            o += flat.size
        assert o == n
        return out

    def upload(self, which: int, channel: int, blocks):
        n = self.chi_elems if which == CHI else self.vertex_elems
        self.upload_packed(which, channel, self._pack(blocks, n))

    def upload_packed(self, which: int, channel: int, packed: np.ndarray):
        """Upload an already packed array (q after q, each block column-major)."""
        n = self.chi_elems if which == CHI else self.vertex_elems
        packed = np.ascontiguousarray(packed, dtype=np.complex128)
        if packed.size != n:
            raise ValueError(f"packed array holds {packed.size} elements, the batch needs {n}")
        _check(self.lib.tups_batch_upload(self._h, which, channel, _dptr(packed)), "tups_batch_upload")

    def download(self, which: int, channel: int):
        n = self.chi_elems if which == CHI else self.vertex_elems
        packed = np.empty(n, dtype=np.complex128)
        _check(self.lib.tups_batch_download(self._h, which, channel, _dptr(packed)), "tups_batch_download")
        out, o = [], 0
        for f, l in zip(self.nf, self.nl):
            shape = (l, l, f) if which == CHI else (l * f, l * f)
            size = int(np.prod(shape))
            out.append(packed[o:o + size].reshape(shape, order="F"))
            o += size
        return out

    def BSE(self, timed: bool = False):
        """The damped ladder of every (q, channel) in one persistent launch."""
        ms = C.c_float()
        _check(self.lib.tups_batch_bse(self._h, C.byref(ms) if timed else None), "tups_batch_bse")
        return float(ms.value) if timed else None

    def BSE_split(self, moduli: int = 12, timed: bool = False):
        """The same update in the split (INT8 CRT) form, class by class of (Nf_q, Nl_q)."""
        ms = C.c_float()
        _check(self.lib.tups_batch_bse_split(self._h, moduli, C.byref(ms) if timed else None),
               "tups_batch_bse_split")
        return float(ms.value) if timed else None

    def sync(self):
        _check(self.lib.tups_batch_sync(self._h), "tups_batch_sync")

    def autotune(self):
        """Pick the fastest batch variant on the resident inputs; returns (name, ms per launch)."""
        v = C.c_int32()
        ms = C.c_float()
        _check(self.lib.tups_batch_autotune(self._h, C.byref(v), C.byref(ms)), "tups_batch_autotune")
        return self.lib.tups_variant_name(K_BSE_BATCH, v.value).decode(), float(ms.value)

    def variants(self):
        return [(v, self.lib.tups_variant_name(K_BSE_BATCH, v).decode())
                for v in range(self.lib.tups_variant_count(K_BSE_BATCH))]

    def set_variant(self, variant: int):
        _check(self.lib.tups_batch_set_variant(self._h, variant), "tups_batch_set_variant")

    def set_product_forms(self, allow_three: bool):
        """Let the batch autotuner pick the three-multiplication complex products."""
        _check(self.lib.tups_batch_set_product_forms(self._h, 1 if allow_three else 0),
               "tups_batch_set_product_forms")


class TupsUneven:
    """Uneven-box parquet iteration: IBZ q = (m, w), q = m Nw + w, with the box
    nf_w[w] per bosonic frequency and nl_m[m] form factors per IBZ momentum; no
    reference counterpart.  `batch` (a TupsBatch view) holds the per-q vertices
    and bubbles, `ctx` (a TupsContext view of the largest box) Lambda; both are
    owned by this object."""

    def __init__(self, nx: int, nf_max: int, nf_w, nl_m, beta: float = 2.0, f_damping: float = 0.5):
        self.lib = load_library()
        self.nf_w = [int(x) for x in nf_w]
        self.nl_m = [int(x) for x in nl_m]
        self.nx, self.nf_max, self.nw = nx, nf_max, nf_max // 2
        h = _V()
        nfa = np.ascontiguousarray(self.nf_w, dtype=np.int32)
        nla = np.ascontiguousarray(self.nl_m, dtype=np.int32)
        _check(self.lib.tups_uneven_create(nx, nf_max, _iptr(nfa), _iptr(nla), beta, f_damping, C.byref(h)),
               "tups_uneven_create")
        self._h = h
        hb, hc = _V(), _V()
        _check(self.lib.tups_uneven_parts(self._h, C.byref(hb), C.byref(hc)), "tups_uneven_parts")
        nibz = len(self.nl_m)
        nf_q = [self.nf_w[q % self.nw] for q in range(nibz * self.nw)]
        nl_q = [self.nl_m[q // self.nw] for q in range(nibz * self.nw)]
        self.batch = TupsBatch(nx, nf_q, nl_q, beta, f_damping, _handle=hb)
        self.ctx = TupsContext(nx, nf_max, max(self.nl_m), beta=beta, f_damping=f_damping, _handle=hc)

    def iteration(self, graph: bool = True):
        """BSE -> crossing (largest box, embedded and restricted) -> G = F - G."""
        _check(self.lib.tups_uneven_iteration(self._h, 1 if graph else 0), "tups_uneven_iteration")

    def sync(self):
        _check(self.lib.tups_uneven_sync(self._h), "tups_uneven_sync")

    def close(self):
        if getattr(self, "_h", None):
            self.batch.close()
            self.ctx.close()
            self.lib.tups_uneven_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
