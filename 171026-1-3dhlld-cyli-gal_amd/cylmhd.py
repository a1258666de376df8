"""ctypes binding of libcylmhd.so (the C ABI declared in include/cylmhd.h).

The library is the product: hand-written gfx950 HIP kernels behind a C ABI.
This module only marshals numpy arrays across that ABI; it never computes
the step itself and it refuses to run when the library is missing, so a test
or benchmark can never pass on a silent fallback.

Arrays use the reference layout A[m][i][j][k] (m = 9 components, k fastest,
ghosts included), FP64, C-contiguous.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcylmhd.so")

LIMITER_MINMOD, LIMITER_MC, LIMITER_VANLEER, LIMITER_SUPERBEE = 0, 1, 2, 3
SOLVER_HLLD, SOLVER_HLL = 0, 1
COOL_NONE, COOL_RAYMOND, COOL_POWERLAW = 0, 1, 2
SOURCE_SPIRAL_COOLING, SOURCE_AXI_COOLING, SOURCE_AXI, SOURCE_UNIFORM_GZ, SOURCE_NONE = 0, 1, 2, 3, 4
GRAV_GALACTIC, GRAV_POINTMASS = 0, 1
PATH_PLAIN64, PATH_FUSED64, PATH_FUSED32 = 0, 1, 2
# slice quantities beyond the primitives 0..8 and slice planes (include/cylmhd.h)
Q_RHO, Q_VR, Q_VPHI, Q_VZ, Q_BR, Q_BPHI, Q_BZ, Q_PR, Q_PSI = range(9)
Q_BENERGY, Q_BETA, Q_TEMP, Q_ONEOTCOOL, Q_VDOTB, Q_DIVB, Q_RHOBZ, Q_SPIRALPOT, Q_VPARA, Q_VPERP = range(9, 19)
NQ = 19
(PLANE_RPHI, PLANE_RPHI_ZMEAN, PLANE_RPHI_ZINT, PLANE_RZ, PLANE_ZR, PLANE_XZ,
 PLANE_XZ_PHIMEAN, PLANE_PHI) = range(8)
OK, EINVAL, ENOMEM, EHIP, EIO, EUNSUP = 0, -1, -2, -3, -4, -5
PATH_NAMES = {PATH_PLAIN64: "plain64", PATH_FUSED64: "fused64", PATH_FUSED32: "fused32"}


class Params(C.Structure):
    """cylmhd_params (include/cylmhd.h); field order must match the header."""
    _fields_ = [
        ("ix", C.c_int), ("jy", C.c_int), ("kz", C.c_int),
        ("lx", C.c_double), ("lz", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("dx0", C.c_double), ("dy0", C.c_double), ("dz0", C.c_double),
        ("stretchratio", C.c_double), ("dxmax", C.c_double), ("dzmax", C.c_double),
        ("rin", C.c_double),
        ("cfl", C.c_double), ("dtmin", C.c_double),
        ("rhocr", C.c_double), ("prcr", C.c_double), ("betacr", C.c_double),
        ("gm", C.c_double), ("gm1", C.c_double), ("oneogm1", C.c_double),
        ("cr", C.c_double), ("te0", C.c_double),
        ("epssp0", C.c_double), ("tausp", C.c_double), ("scalez", C.c_double),
        ("scaler", C.c_double), ("scalephase", C.c_double),
        ("armnum", C.c_int),
        ("omegasp", C.c_double), ("omegapattern", C.c_double), ("pitch", C.c_double),
        ("limiter", C.c_int), ("solver", C.c_int), ("rk", C.c_int),
        ("cooling", C.c_int), ("gravity", C.c_int), ("source", C.c_int),
        ("gm_point", C.c_double),
        ("cool_lambda0", C.c_double), ("cool_tmin", C.c_double), ("cool_index", C.c_double),
        ("path", C.c_int),
        ("disc_h", C.c_double), ("disc_floor", C.c_double), ("disc_cs2", C.c_double),
        ("disc_beta", C.c_double), ("beta_lo", C.c_double), ("beta_hi", C.c_double),
        ("beta_mode", C.c_int),
        ("pert_amp", C.c_double),
        ("seed", C.c_uint64),
    ]

    def copy(self):
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p

    @property
    def dims(self):
        return self.ix + 2, self.jy + 2, self.kz + 2


class Diag(C.Structure):
    _fields_ = [("mass", C.c_double), ("energy", C.c_double),
                ("divb_l1", C.c_double), ("psi_l1", C.c_double)]


class IdlPlanes(C.Structure):
    _fields_ = [("kplane", C.c_int), ("k100", C.c_int), ("jplane", C.c_int), ("azim_i", C.c_int * 4),
                ("azim_k", C.c_int)]


class CylmhdError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    """Load libcylmhd.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise CylmhdError(f"{LIB_PATH} is missing: run __graft_entry__.build() (there is no fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(Params)
    sig = {
        "cylmhd_params_default": ([P], C.c_int),
        "cylmhd_params_config": ([P, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "cylmhd_grid_host": ([P, _dp, _dp, _dp], C.c_int),
        "cylmhd_model_disc": ([P, _dp], C.c_int),
        "cylmhd_model_galactic_disc3": ([P, _dp], C.c_int),
        "cylmhd_model_disc_state": ([P, _dp, _dp], C.c_int),
        "cylmhd_fill_ghosts_host": ([P, _dp, _dp], C.c_int),
        "cylmhd_vtou_host": ([P, _dp, _dp], C.c_int),
        "cylmhd_check_host": ([P, C.c_void_p, _dp, _dp, C.POINTER(C.c_int)], C.c_int),
        "cylmhd_create": ([P, C.POINTER(C.c_void_p)], C.c_int),
        "cylmhd_destroy": ([C.c_void_p], C.c_int),
        "cylmhd_dims": ([C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "cylmhd_set_initial": ([C.c_void_p, _dp, _dp], C.c_int),
        "cylmhd_upload_state": ([C.c_void_p, _dp, _dp], C.c_int),
        "cylmhd_download_state": ([C.c_void_p, _dp, _dp], C.c_int),
        "cylmhd_setdt": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_int)], C.c_int),
        "cylmhd_setdt_noalloc": ([P, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)], C.c_int),
        "cylmhd_gravity": ([C.c_void_p, C.c_double], C.c_int),
        "cylmhd_step": ([C.c_void_p, C.c_double, C.c_double], C.c_int),
        "cylmhd_floor_counts": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
        "cylmhd_diagnostics": ([C.c_void_p, C.POINTER(Diag)], C.c_int),
        "cylmhd_sync": ([C.c_void_p], C.c_int),
        "cylmhd_kernel_timing": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int)], C.c_int),
        "cylmhd_error_string": ([C.c_int], C.c_char_p),
        "cylmhd_binary_output": ([C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                  _dp, _dp, _dp, _dp], C.c_int),
        "cylmhd_restart_read": ([C.c_char_p, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
        "cylmhd_slice_size": ([C.c_void_p, C.c_int, C.POINTER(C.c_size_t)], C.c_int),
        "cylmhd_slice": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_size_t], C.c_int),
        "cylmhd_gdl_output": ([C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
        "cylmhd_gdl_output_ctx": ([C.c_void_p, C.c_char_p, C.c_int], C.c_int),
        "cylmhd_idl_output": ([C.c_void_p, C.c_char_p, C.c_int, C.POINTER(IdlPlanes)], C.c_int),
        "cylmhd_params_batch": ([P, C.c_int, C.c_uint64], C.c_int),
        "cylmhd_batch_create": ([P, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
        "cylmhd_batch_destroy": ([C.c_void_p], C.c_int),
        "cylmhd_batch_disc": ([C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
        "cylmhd_batch_tune": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "cylmhd_batch_run": ([C.c_void_p, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "cylmhd_batch_disc_tile": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "cylmhd_batch_status": ([C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "cylmhd_batch_floor_counts": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    }
    for name, (args, res) in sig.items():
        try:
            fn = getattr(L, name)
        except AttributeError:
            continue          # reported by missing_symbols(); calling it raises
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def missing_symbols():
    L = lib()
    return [s for s in exported_symbols() if not hasattr(L, s)]


def check(rc, what=""):
    if rc != 0:
        msg = lib().cylmhd_error_string(rc).decode()
        raise CylmhdError(f"{what}: {msg} ({rc})", rc)


def params_default():
    p = Params()
    check(lib().cylmhd_params_default(C.byref(p)), "params_default")
    return p


def params_config(config, nr=0, nphi=0, nz=0):
    p = Params()
    check(lib().cylmhd_params_config(C.byref(p), config, nr, nphi, nz), "params_config")
    return p


def grid(p):
    NX, NY, NZ = p.dims
    x, y, z = np.zeros(NX), np.zeros(NY), np.zeros(NZ)
    check(lib().cylmhd_grid_host(C.byref(p), x, y, z), "grid")
    return x, y, z


def model_disc(p):
    NX, NY, NZ = p.dims
    V = np.zeros((9, NX, NY, NZ))
    check(lib().cylmhd_model_disc(C.byref(p), V), "model_disc")
    return V


def disc_state(p):
    """Consistent initial (U, V) of the synthetic disc (ghosts filled)."""
    NX, NY, NZ = p.dims
    U = np.zeros((9, NX, NY, NZ))
    V = np.zeros((9, NX, NY, NZ))
    check(lib().cylmhd_model_disc_state(C.byref(p), U, V), "model_disc_state")
    return U, V


def fill_ghosts(p, U, V):
    """Ghosts of (U, V) in place by the RK3 cell-boundary set."""
    check(lib().cylmhd_fill_ghosts_host(C.byref(p), U, V), "fill_ghosts")
    return U, V


def model_galactic_disc3(p):
    NX, NY, NZ = p.dims
    V = np.zeros((9, NX, NY, NZ))
    check(lib().cylmhd_model_galactic_disc3(C.byref(p), V), "model_galactic_disc3")
    return V


def vtou(p, V):
    U = np.zeros_like(V)
    check(lib().cylmhd_vtou_host(C.byref(p), np.ascontiguousarray(V), U), "vtou")
    return U


def check_host(p, U, V, rho_init=None):
    """Check (check.h:23-188) on host arrays U, V in place, as RestartRead
    runs it; rho_init = the SaveInitial density (floors rho_init * rhocr) or
    None (no SaveInitial yet: no density floor).  Returns the three counts."""
    for a in (U, V):
        if not (a.dtype == np.float64 and a.flags.c_contiguous):
            raise ValueError("check_host works in place on C-contiguous FP64 arrays")
    ri = None
    if rho_init is not None:
        ri = np.ascontiguousarray(rho_init, dtype=np.float64)
        if ri.size != U[0].size:
            raise ValueError("rho_init must have one value per cell")
    c = (C.c_int * 3)()
    check(lib().cylmhd_check_host(C.byref(p), None if ri is None else ri.ctypes.data_as(C.c_void_p), U, V, c),
          "check_host")
    return [c[0], c[1], c[2]]


def setdt_noalloc(p, V_dev, minlength_dev, minminlength, stream=None):
    """cylmhd_setdt_noalloc on device arrays given by address (V: [9][NX][NY][NZ]
    FP64 primitives, minlength: [NX][NZ] FP64); returns (dt, ch, (i, j, k))."""
    dt, ch = C.c_double(), C.c_double()
    ijk = (C.c_int * 3)()
    check(lib().cylmhd_setdt_noalloc(C.byref(p), C.c_void_p(V_dev), C.c_void_p(minlength_dev), minminlength,
                                     C.c_void_p(stream) if stream else None, C.byref(dt), C.byref(ch), ijk),
          "setdt_noalloc")
    return dt.value, ch.value, tuple(ijk)


def binary_output(dirname, t, x, y, z, V):
    nx, ny, nz = V.shape[1:]
    check(lib().cylmhd_binary_output(dirname.encode(), nx, ny, nz, t, x, y, z,
                                     np.ascontiguousarray(V)), "binary_output")


def gdl_output(dirname, nout, x, y, z, V):
    """GDLOutputCylinder's text files (output.h:280-384) from host arrays."""
    nx, ny, nz = V.shape[1:]
    check(lib().cylmhd_gdl_output(dirname.encode(), nout, nx, ny, nz, x, y, z,
                                  np.ascontiguousarray(V)), "gdl_output")


def restart_read(dirname, shape):
    V = np.zeros((9,) + tuple(shape))
    check(lib().cylmhd_restart_read(dirname.encode(), shape[0], shape[1], shape[2], V), "restart_read")
    return V


class Context:
    """One device context (one HIP stream).  Mirrors the reference call
    sequence: set_initial (SaveInitial), setdt (Setdt), gravity
    (SpiralGravity), step (TVDRK)."""

    def __init__(self, p):
        self.p = p.copy()
        self.dims = p.dims
        h = C.c_void_p()
        check(lib().cylmhd_create(C.byref(self.p), C.byref(h)), "create")
        self.h = h

    def close(self):
        if self.h:
            lib().cylmhd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_initial(self, U, V):
        check(lib().cylmhd_set_initial(self.h, np.ascontiguousarray(U), np.ascontiguousarray(V)), "set_initial")

    def upload(self, U, V):
        check(lib().cylmhd_upload_state(self.h, np.ascontiguousarray(U), np.ascontiguousarray(V)), "upload")

    def download(self):
        U = np.zeros((9,) + self.dims)
        V = np.zeros((9,) + self.dims)
        check(lib().cylmhd_download_state(self.h, U, V), "download")
        return U, V

    def slice(self, quantity, plane, a=0, b=0):
        """One plane of a primitive or derived quantity, cut on the device."""
        n = C.c_size_t()
        check(lib().cylmhd_slice_size(self.h, plane, C.byref(n)), "slice_size")
        out = np.zeros(n.value)
        check(lib().cylmhd_slice(self.h, quantity, plane, a, b, out, n.value), "slice")
        return out

    def gdl_output(self, dirname, nout):
        check(lib().cylmhd_gdl_output_ctx(self.h, dirname.encode(), nout), "gdl_output_ctx")

    def idl_output(self, root, nout, kplane=1, k100=100, jplane=1, azim_i=(0, 0, 0, 0), azim_k=0):
        """The text slices the reference's IDL scripts read (idl<nout>/...,
        azimuth-K/...)."""
        pl = IdlPlanes(kplane, k100, jplane, (C.c_int * 4)(*azim_i), azim_k)
        check(lib().cylmhd_idl_output(self.h, root.encode(), nout, C.byref(pl)), "idl_output")

    def setdt(self):
        dt, ch = C.c_double(), C.c_double()
        ijk = (C.c_int * 3)()
        check(lib().cylmhd_setdt(self.h, C.byref(dt), C.byref(ch), ijk), "setdt")
        return dt.value, ch.value, tuple(ijk)

    def gravity(self, t):
        check(lib().cylmhd_gravity(self.h, t), "gravity")

    def step(self, dt, ch):
        check(lib().cylmhd_step(self.h, dt, ch), "step")

    def sync(self):
        check(lib().cylmhd_sync(self.h), "sync")

    def floor_counts(self):
        c = (C.c_int * 9)()
        check(lib().cylmhd_floor_counts(self.h, c), "floor_counts")
        return np.array(list(c)).reshape(3, 3)

    def diagnostics(self):
        d = Diag()
        check(lib().cylmhd_diagnostics(self.h, C.byref(d)), "diagnostics")
        return dict(mass=d.mass, energy=d.energy, divb_l1=d.divb_l1, psi_l1=d.psi_l1)

    def kernel_timing(self, which=0, reset=False):
        ms, n = C.c_double(), C.c_int()
        check(lib().cylmhd_kernel_timing(self.h, which, int(reset), C.byref(ms), C.byref(n)), "kernel_timing")
        return ms.value, n.value

    def run(self, nsteps, t=0.0):
        """main.c:316-393 loop (no dumps): returns (t, [dt...])."""
        dts = []
        for _ in range(nsteps):
            dt, ch, _ = self.setdt()
            if not np.isfinite(dt) or dt < self.p.dtmin:
                raise CylmhdError(f"stop due to small dt or inf (dt={dt})")
            self.gravity(t)
            self.step(dt, ch)
            t += dt
            dts.append(dt)
        return t, dts


def exported_symbols():
    """Names declared in include/cylmhd.h (for the export test)."""
    import re
    hdr = os.path.join(HERE, "..", "include", "cylmhd.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(cylmhd_[a-z0-9_]+)\s*\(", txt)))


def params_batch(n, seed=20171026):
    """Parameters of the benchmark batch (cylmhd_params_batch)."""
    arr = (Params * n)()
    check(lib().cylmhd_params_batch(arr, n, seed), "params_batch")
    return [arr[d].copy() for d in range(n)]


class _DiscRef(Context):
    """A disc context owned by a Batch (not destroyed from Python)."""

    def __init__(self, p, h):
        self.p = p.copy()
        self.dims = p.dims
        self.h = h

    def close(self):
        self.h = None


class Batch:
    """Many discs advanced together (cylmhd_batch_*): one work list, one
    launch per RK stage, per-disc dt on the device, the step replayed from a
    captured hipGraph."""

    def __init__(self, params):
        self.params = [p.copy() for p in params]
        n = len(params)
        arr = (Params * n)(*self.params)
        h = C.c_void_p()
        check(lib().cylmhd_batch_create(arr, n, C.byref(h)), "batch_create")
        self.h = h
        self.n = n

    def disc(self, d):
        h = C.c_void_p()
        check(lib().cylmhd_batch_disc(self.h, d, C.byref(h)), "batch_disc")
        return _DiscRef(self.params[d], h)

    def tune(self, steps=2):
        ch = (C.c_int * 4)()
        check(lib().cylmhd_batch_tune(self.h, steps, ch), "batch_tune")
        return tuple(ch)

    def disc_tile(self, d):
        """(shape, phi cells, z cells, threads) of the tile disc d runs with."""
        t = (C.c_int * 4)()
        check(lib().cylmhd_batch_disc_tile(self.h, d, t), "batch_disc_tile")
        return tuple(t)

    def run(self, nsteps):
        ms = C.c_double()
        check(lib().cylmhd_batch_run(self.h, nsteps, C.byref(ms)), "batch_run")
        return ms.value

    def status(self, d):
        t, dt = C.c_double(), C.c_double()
        st, al = C.c_int(), C.c_int()
        check(lib().cylmhd_batch_status(self.h, d, C.byref(t), C.byref(dt), C.byref(st), C.byref(al)),
              "batch_status")
        return dict(t=t.value, dt=dt.value, steps=st.value, alive=bool(al.value))

    def floor_counts(self, d):
        c = (C.c_int * 9)()
        check(lib().cylmhd_batch_floor_counts(self.h, d, c), "batch_floor_counts")
        return np.array(list(c)).reshape(3, 3)

    def close(self):
        if self.h:
            lib().cylmhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
