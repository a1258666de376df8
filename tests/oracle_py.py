"""ctypes binding of oracle/liboracle.so -- the CPU FP64 restatement.

Test infrastructure: only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _load(params_cls):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(ORACLE_LIB):
        raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
    L = C.CDLL(ORACLE_LIB)
    P = C.POINTER(params_cls)
    vp = C.c_void_p
    sig = {
        "orc_create": ([P], vp),
        "orc_destroy": ([vp], None),
        "orc_dims": ([vp, C.POINTER(C.c_int)] * 1 + [C.POINTER(C.c_int), C.POINTER(C.c_int)], None),
        "orc_grid": ([vp, _dp, _dp, _dp, _dp, C.POINTER(C.c_double)], None),
        "orc_geometry": ([vp, _dp], None),
        "orc_dzm": ([vp, _dp], None),
        "orc_absorb_factors": ([vp, _dp, _dp, _dp], C.c_int),
        "orc_gravity_axi": ([vp, _dp, _dp], None),
        "orc_set_initial": ([vp, _dp, _dp], None),
        "orc_set_state": ([vp, _dp, _dp], None),
        "orc_restart_state": ([vp, _dp, C.POINTER(C.c_int)], None),
        "orc_get_state": ([vp, _dp, _dp], None),
        "orc_setdt": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)], None),
        "orc_gravity": ([vp, C.c_double], None),
        "orc_get_gravity": ([vp, _dp, _dp, _dp], None),
        "orc_step": ([vp, C.c_double, C.c_double], None),
        "orc_stage": ([vp, C.c_int, C.c_double, C.c_double], None),
        "orc_get_fluxes": ([vp, _dp], None),
        "orc_floor_counts": ([vp, C.POINTER(C.c_int)], None),
        "orc_diagnostics": ([vp, C.c_void_p], None),
        "orc_run": ([vp, C.c_int, C.c_double, C.POINTER(C.c_double)], C.c_double),
        "orc_threads": ([], C.c_int),
        "orc_cooling": ([P, C.c_double, C.c_double], C.c_double),
        "orc_hlld_face": ([vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp], None),
        "orc_hll_face": ([vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


class Oracle:
    def __init__(self, p):
        self.P = type(p)
        self.L = _load(self.P)
        self.p = p.copy()
        self.h = self.L.orc_create(C.byref(self.p))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.dims = p.dims
        NX, NY, NZ = self.dims
        self.N = NX * NY * NZ
        self.NF = (NX - 1) * (NY - 1) * (NZ - 1)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        NX, NY, NZ = self.dims
        x, xm, y, z = np.zeros(NX), np.zeros(NX), np.zeros(NY), np.zeros(NZ)
        mml = C.c_double()
        self.L.orc_grid(self.h, x, xm, y, z, C.byref(mml))
        return dict(x=x, xm=xm, y=y, z=z, minminlength=mml.value)

    def geometry(self):
        NX, NY, NZ = self.dims
        out = np.zeros((6, NX, NZ))
        self.L.orc_geometry(self.h, out)
        return out

    def dzm(self):
        out = np.zeros(self.dims[2])
        self.L.orc_dzm(self.h, out)
        return out

    def absorb_factors(self):
        NX, NY, NZ = self.dims
        xabs, zabs = np.zeros(11), np.zeros(11)
        cabs = np.zeros((NX + 1) * NZ)
        irin = self.L.orc_absorb_factors(self.h, xabs, zabs, cabs)
        return xabs, zabs, cabs[: (irin + 1) * NZ].reshape(irin + 1, NZ), irin

    def gravity_axi(self):
        NX, NY, NZ = self.dims
        gx, gz = np.zeros((NX, NZ)), np.zeros((NX, NZ))
        self.L.orc_gravity_axi(self.h, gx, gz)
        return gx, gz

    def set_initial(self, U, V):
        self.L.orc_set_initial(self.h, np.ascontiguousarray(U), np.ascontiguousarray(V))

    def set_state(self, U, V):
        self.L.orc_set_state(self.h, np.ascontiguousarray(U), np.ascontiguousarray(V))

    def restart_state(self, V):
        """RestartRead's tail (restartread.h:132-139): VtoU of V, then Check
        with the current floors; returns the three Check counts."""
        c = (C.c_int * 3)()
        self.L.orc_restart_state(self.h, np.ascontiguousarray(V, dtype=np.float64), c)
        return [c[0], c[1], c[2]]

    def state(self):
        U = np.zeros((9,) + self.dims)
        V = np.zeros((9,) + self.dims)
        self.L.orc_get_state(self.h, U, V)
        return U, V

    def setdt(self):
        dt, ch = C.c_double(), C.c_double()
        ijk = (C.c_int * 3)()
        self.L.orc_setdt(self.h, C.byref(dt), C.byref(ch), ijk)
        return dt.value, ch.value, tuple(ijk)

    def gravity(self, t):
        self.L.orc_gravity(self.h, t)

    def get_gravity(self):
        g = [np.zeros(self.dims) for _ in range(3)]
        self.L.orc_get_gravity(self.h, *g)
        return g

    def step(self, dt, ch):
        self.L.orc_step(self.h, dt, ch)

    def stage(self, s, dt, ch):
        self.L.orc_stage(self.h, s, dt, ch)

    def fluxes(self):
        NX, NY, NZ = self.dims
        out = np.zeros((3, 9, NX - 1, NY - 1, NZ - 1))
        self.L.orc_get_fluxes(self.h, out)
        return out

    def floor_counts(self):
        c = (C.c_int * 9)()
        self.L.orc_floor_counts(self.h, c)
        return np.array(list(c)).reshape(3, 3)

    def diagnostics(self):
        d = (C.c_double * 4)()
        self.L.orc_diagnostics(self.h, C.cast(d, C.c_void_p))
        return dict(mass=d[0], energy=d[1], divb_l1=d[2], psi_l1=d[3])

    def run(self, nsteps, t=0.0):
        el = C.c_double()
        t = self.L.orc_run(self.h, nsteps, t, C.byref(el))
        return t, el.value

    def run_steps(self, nsteps, t=0.0):
        """main.c loop returning per-step (dt, ch, ijk)."""
        hist = []
        for _ in range(nsteps):
            dt, ch, ijk = self.setdt()
            if not np.isfinite(dt) or dt < self.p.dtmin:
                break
            self.gravity(t)
            self.step(dt, ch)
            t += dt
            hist.append((dt, ch, ijk))
        return t, hist

    def hlld_face(self, d, i, ch, vl, vr, hll=False):
        out = np.zeros(9)
        fn = self.L.orc_hll_face if hll else self.L.orc_hlld_face
        fn(self.h, d, i, ch, np.ascontiguousarray(vl, dtype=np.float64),
           np.ascontiguousarray(vr, dtype=np.float64), out)
        return out


def ref_harness(name):
    path = os.path.join(REF_DIR, name, "ref_harness")
    return path if os.path.isfile(path) else None


def run_ref(name, V_init, V_cur, grx, grz, nsteps, rk, galactic=False, dumpdir=None, U_init=None, U_cur=None,
            model=False):
    """Run one compiled reference harness; returns (dict of arrays, stdout).
    With U_init/U_cur the harness takes the conserved states as given
    instead of rebuilding them with VtoU; with model=True it starts from the
    reference's own GalacticDisc3 (the input states only set the sizes)."""
    exe = ref_harness(name)
    if exe is None:
        raise FileNotFoundError(name)
    _, NX, NY, NZ = V_init.shape
    N = NX * NY * NZ
    NF = (NX - 1) * (NY - 1) * (NZ - 1)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            flags = (1 if galactic else 0) | (2 if U_init is not None else 0) | (4 if model else 0)
            np.array([NX, NY, NZ, flags], dtype=np.int64).tofile(f)
            np.ascontiguousarray(V_init, dtype=np.float64).tofile(f)
            np.ascontiguousarray(V_cur, dtype=np.float64).tofile(f)
            np.ascontiguousarray(grx, dtype=np.float64).tofile(f)
            np.ascontiguousarray(grz, dtype=np.float64).tofile(f)
            if U_init is not None:
                np.ascontiguousarray(U_init, dtype=np.float64).tofile(f)
                np.ascontiguousarray(U_cur, dtype=np.float64).tofile(f)
        args = [exe, fin, fout, str(nsteps), str(rk)]
        if dumpdir:
            args.append(dumpdir)
        r = subprocess.run(args, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        if r.returncode != 0:
            raise RuntimeError(f"{name} failed: {r.stderr}")
        raw = np.fromfile(fout, dtype=np.uint8)
    hdr = raw[:32].view(np.int64)
    n = int(hdr[3])
    d = raw[32:].view(np.float64)
    o = 0

    def take(cnt, shape=None):
        nonlocal o
        a = d[o:o + cnt].copy()
        o += cnt
        return a.reshape(shape) if shape else a

    out = {}
    out["dt"] = take(n)
    out["ch"] = take(n)
    out["U"] = take(9 * N, (9, NX, NY, NZ))
    out["V"] = take(9 * N, (9, NX, NY, NZ))
    out["F"] = take(9 * NF, (9, NX - 1, NY - 1, NZ - 1))
    out["G"] = take(9 * NF, (9, NX - 1, NY - 1, NZ - 1))
    out["H"] = take(9 * NF, (9, NX - 1, NY - 1, NZ - 1))
    out["x"] = take(NX)
    out["xm"] = take(NX)
    out["y"] = take(NY)
    out["z"] = take(NZ)
    out["minminlength"] = take(1)[0]
    out["geom"] = take(6 * NX * NZ, (6, NX, NZ))
    out["grx"] = take(NX * NZ, (NX, NZ))
    out["grz"] = take(NX * NZ, (NX, NZ))
    out["xabs"] = take(11)
    out["cabs"] = d[o:].copy()
    return out, r.stdout


def parse_counts(stdout):
    """Check() prints three counters per stage (check.h:180-183)."""
    vals = []
    for line in stdout.splitlines():
        for key in ("negative density point :", "negative pressure point :", "critical beta point :"):
            if line.startswith(key):
                vals.append(int(line.split(":")[1]))
    return np.array(vals).reshape(-1, 3)


def parse_dt_mesh(stdout):
    out = []
    for line in stdout.splitlines():
        if line.startswith("dt factor mesh"):
            out.append(tuple(int(v) for v in line.split()[3:6]))
    return out
