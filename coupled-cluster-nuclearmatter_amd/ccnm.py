"""Python mirror of the engine's C ABI (include/ccgpu.h) over ctypes.

The names follow the reference's hot-path routines (ccd_iter, t2_intermediate,
diis_t1t2, ccd_energy_save, setup_t_amplitudes, pert_triples2) so host code and
tests read like the reference's own driver.  All compute runs in the HIP library
libccgpu.so next to this file; if that library is missing every call fails with
an ImportError/RuntimeError: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccgpu.so")

PNM, SNM = 0, 1
INT_MINNESOTA = 0
INT_CHIRAL = 1
LO, NLO, NNLO = 0, 2, 3
BLOCK_HHHH, BLOCK_HHHP, BLOCK_HHPP, BLOCK_HPHP, BLOCK_HPPP, BLOCK_PPPP = 1, 2, 3, 4, 5, 6
PREC_FP64, PREC_FP32, PREC_V32, PREC_BF16X3 = 0, 1, 2, 3
ACCEL_REFERENCE, ACCEL_ANDERSON = 0, 1
STATUS_DIIS_FALLBACK, STATUS_NONFINITE, STATUS_MAXIT = 1, 2, 4


class LatticeParams(C.Structure):
    _fields_ = [("A", C.c_int), ("matter", C.c_int), ("rho", C.c_double), ("nmax", C.c_int),
                ("twist", C.c_double * 3)]


class Dims(C.Structure):
    _fields_ = [("nsp", C.c_int), ("nholes", C.c_int), ("nparticles", C.c_int),
                ("nchannels", C.c_int), ("nph_channels", C.c_int), ("max_npp", C.c_int),
                ("max_nhh", C.c_int), ("t2_unique", C.c_longlong), ("t2_ordered", C.c_longlong),
                ("vpppp_unique", C.c_longlong), ("ring_elems", C.c_longlong),
                ("box", C.c_double), ("volume", C.c_double), ("kf", C.c_double),
                ("pp_ladder_flops", C.c_double), ("hh_ladder_flops", C.c_double), ("ring_flops", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("subspace", C.c_int), ("diis_step", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_int), ("precision", C.c_int), ("refine_every", C.c_int),
                ("use_graph", C.c_int), ("verbose", C.c_int), ("fp32_switch", C.c_double),
                ("ccdt1", C.c_int), ("accel", C.c_int), ("mix", C.c_double), ("tail", C.c_int),
                ("confirm", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("e0", C.c_double * 2),
                ("embpt2", C.c_double * 2), ("eccd", C.c_double * 2),
                ("ecorr_last", C.c_double * 2), ("seconds", C.c_double),
                ("fp32_iterations", C.c_int), ("status", C.c_int)]

    def as_dict(self):
        z = lambda a: complex(a[0], a[1])
        return {"iterations": self.iterations, "converged": bool(self.converged),
                "e0": z(self.e0), "embpt2": z(self.embpt2), "eccd": z(self.eccd),
                "ecorr": z(self.ecorr_last), "seconds": self.seconds,
                "fp32_iterations": self.fp32_iterations, "status": self.status}


# every exported symbol with its argument types (kept in step with include/ccgpu.h)
_I, _D, _P = C.c_int, C.c_double, C.c_void_p
_IP, _DP, _FP = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
SIGNATURES = {
    "ccgpu_plan": [C.POINTER(LatticeParams), C.POINTER(Dims)],
    "ccgpu_create": [C.POINTER(LatticeParams), _IP],
    "ccgpu_create_from_orbitals": [_I, _I, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _D, _D, _IP],
    "ccgpu_destroy": [_I],
    "ccgpu_get_dims": [_I, C.POINTER(Dims)],
    "ccgpu_get_orbitals": [_I, _IP, _IP, _IP, _IP, _IP, _DP],
    "ccgpu_generate_interaction": [_I, _I],
    "ccgpu_set_chiral": [_I, _I, _D, _DP],
    "ccgpu_vnn_elements": [_I, _I, _IP, _IP, _IP, _IP, _DP],
    "ccgpu_set_3nf": [_I, _I, _D, _D, _D, _I],
    "ccgpu_set_block": [_I, _I, _I, _IP, _IP, _I, _IP, _IP, _DP, _I],
    "ccgpu_set_fock": [_I, _DP, _D, _D],
    "ccgpu_finalize_setup": [_I],
    "ccgpu_get_e0": [_I, _DP],
    "ccgpu_get_fock_diag": [_I, _DP],
    "ccgpu_mbpt2_start": [_I, _DP],
    "ccgpu_residual": [_I, _I],
    "ccgpu_update": [_I, _IP, _IP, _I, _I],
    "ccgpu_energy": [_I, _DP],
    "ccgpu_solve": [_I, C.POINTER(SolveOpts), C.POINTER(SolveResult)],
    "ccgpu_triples": [_I, _DP],
    "ccgpu_triples_parts": [_I, _DP, _DP],
    "ccgpu_setup_t3": [_I, _DP],
    "ccgpu_t3_intermediate": [_I],
    "ccgpu_t3_energy": [_I, _DP],
    "ccgpu_triples_stats": [_I, C.POINTER(C.c_longlong), _DP],
    "ccgpu_get_t2": [_I, _I, _IP, _IP, _IP, _IP, _DP],
    "ccgpu_set_t2": [_I, _I, _IP, _IP, _IP, _IP, _DP],
    "ccgpu_get_t2_unique": [_I, _DP],
    "ccgpu_set_t2_unique": [_I, _DP],
    "ccgpu_get_t2_labels": [_I, _IP],
    "ccgpu_batch_solve": [_I, _IP, C.POINTER(SolveOpts), C.POINTER(SolveResult)],
    "ccgpu_solve_energies": [_I, _DP, _I, _IP],
    "ccgpu_kernel_stats": [C.c_char_p, _I, _DP, C.POINTER(C.c_longlong), _I],
    "ccgpu_debug_fetch": [_I, _I, _DP, C.c_longlong],
    "ccgpu_mlp_create": [C.c_ulonglong, _IP],
    "ccgpu_mlp_destroy": [_I],
    "ccgpu_mlp_get_params": [_I, _FP, _I],
    "ccgpu_mlp_set_params": [_I, _FP, _I],
    "ccgpu_mlp_add_solve": [_I, _I, C.POINTER(C.c_longlong)],
    "ccgpu_mlp_train": [_I, _I, _I, _D, C.c_ulonglong, _DP],
    "ccgpu_mlp_warm_start": [_I, _I, _DP],
    "ccgpu_mlp_warm_start_batch": [_I, _IP, _I],
    "ccgpu_hbm_read_probe": [C.c_longlong, _I, _DP],
    "ccgpu_mlp_forward": [_I, _I, _FP, _FP],
    "ccgpu_mlp_loss_grad": [_I, _I, _FP, _FP, _DP, _FP],
    "ccgpu_mlp_features": [_I, _I, _FP, _FP],
    "ccgpu_autotune_report": [C.c_char_p, _I],
    "ccgpu_energy_noalloc": [C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "ccgpu_last_error": [C.c_char_p, _I],
}

_lib = None


def lib():
    """Load libccgpu.so (built in-tree); raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        buf = C.create_string_buffer(2048)
        lib().ccgpu_last_error(buf, 2048)
        raise RuntimeError(buf.value.decode())


def _ip(a):
    return a.ctypes.data_as(_IP)


def _dp(a):
    return a.ctypes.data_as(_DP)


def plan(A, matter, rho, nmax, twist=(0.0, 0.0, 0.0)):
    """Host-only lattice plan (no GPU): dimensions of the channel structure."""
    p = LatticeParams(A, matter, rho, nmax, (C.c_double * 3)(*twist))
    d = Dims()
    _check(lib().ccgpu_plan(C.byref(p), C.byref(d)))
    return {k: getattr(d, k) for k, _ in Dims._fields_}


def default_opts(subspace=6, diis_step=6, tol=1e-6, maxit=100, precision=PREC_FP64,
                 refine_every=0, use_graph=0, verbose=0, fp32_switch=1.0, ccdt1=0,
                 accel=ACCEL_REFERENCE, mix=0.0, tail=PREC_FP64, confirm=0):
    return SolveOpts(subspace, diis_step, tol, maxit, precision, refine_every, use_graph, verbose,
                     fp32_switch, int(ccdt1), int(accel), float(mix), int(tail), int(confirm))


class Solver:
    """One lattice solve on the GPU (ccd_iter and its callees)."""

    def __init__(self, A=None, matter=PNM, rho=0.16, nmax=1, twist=(0.0, 0.0, 0.0), orbitals=None):
        h = C.c_int(0)
        if orbitals is None:
            p = LatticeParams(A, matter, rho, nmax, (C.c_double * 3)(*twist))
            _check(lib().ccgpu_create(C.byref(p), C.byref(h)))
        else:
            o = orbitals
            arr_i = [np.ascontiguousarray(o[k], np.int32) for k in ("nx", "ny", "nz", "sz", "tz")]
            arr_d = [np.ascontiguousarray(o[k], np.float64) for k in ("kx", "ky", "kz", "e")]
            nsp = len(arr_d[3])
            _check(lib().ccgpu_create_from_orbitals(nsp, int(o["A"]), *[_ip(a) for a in arr_i],
                                                    *[_dp(a) for a in arr_d], float(o["box"]),
                                                    float(o["volume"]), C.byref(h)))
        self.h = h.value
        self.dims = self.get_dims()

    def close(self):
        if self.h:
            lib().ccgpu_destroy(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_dims(self):
        d = Dims()
        _check(lib().ccgpu_get_dims(self.h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in Dims._fields_}

    def orbitals(self):
        n = self.dims["nsp"]
        out = {k: np.zeros(n, np.int32) for k in ("nx", "ny", "nz", "sz", "tz")}
        e = np.zeros(n)
        _check(lib().ccgpu_get_orbitals(self.h, *[_ip(out[k]) for k in ("nx", "ny", "nz", "sz", "tz")], _dp(e)))
        out["e"] = e
        return out

    # setup
    def set_chiral(self, order=NNLO, cutoff=450.0, ci=None):
        """Chiral order and cutoff of INT_CHIRAL (ccm_in's first line); ci
        replaces the pi-N LECs c1..c4 (GeV^-1)."""
        arr = None if ci is None else np.ascontiguousarray(ci, np.float64)
        if arr is not None and arr.shape != (4,):
            raise ValueError("ci must hold c1, c2, c3, c4")
        _check(lib().ccgpu_set_chiral(self.h, int(order), float(cutoff),
                                      None if arr is None else arr.ctypes.data_as(_DP)))

    def set_3nf(self, approx=2, cE=0.017, cD=0.79, cutoff=450.0, nexp=3):
        """NNLO three-nucleon force of INT_CHIRAL, normal-ordered to zero-,
        one- or two-body (approx 0/1/2; -1 off), as tnf_approx of ccm_in."""
        _check(lib().ccgpu_set_3nf(self.h, int(approx), float(cE), float(cD), float(cutoff), int(nexp)))

    def generate_interaction(self, kind=INT_MINNESOTA):
        _check(lib().ccgpu_generate_interaction(self.h, kind))

    def vnn_elements(self, p, q, r, s):
        """<pq||rs> of the generated interaction (0-based orbital indices)."""
        arrs = [np.ascontiguousarray(np.ravel(x), np.int32) for x in (p, q, r, s)]
        n = len(arrs[0])
        out = np.zeros(n, np.complex128)
        _check(lib().ccgpu_vnn_elements(self.h, n, *[a.ctypes.data_as(_IP) for a in arrs],
                                        out.ctypes.data_as(_DP)))
        return out

    def set_block(self, kind, bra_p, bra_q, ket_r, ket_s, val):
        """Reference-layout block: 1-based pair lists, val[nbra, nket] complex."""
        bp, bq = np.ascontiguousarray(bra_p, np.int32), np.ascontiguousarray(bra_q, np.int32)
        kr, ks = np.ascontiguousarray(ket_r, np.int32), np.ascontiguousarray(ket_s, np.int32)
        v = np.asfortranarray(np.asarray(val, np.complex128))
        _check(lib().ccgpu_set_block(self.h, kind, len(bp), _ip(bp), _ip(bq), len(kr), _ip(kr), _ip(ks),
                                     v.ctypes.data_as(_DP), len(bp)))

    def set_fock(self, fock, e0):
        f = np.asfortranarray(np.asarray(fock, np.complex128))
        _check(lib().ccgpu_set_fock(self.h, f.ctypes.data_as(_DP), float(np.real(e0)), float(np.imag(e0))))

    def finalize_setup(self):
        _check(lib().ccgpu_finalize_setup(self.h))

    def e0(self):
        x = np.zeros(2)
        _check(lib().ccgpu_get_e0(self.h, _dp(x)))
        return complex(x[0], x[1])

    def fock_diag(self):
        x = np.zeros(2 * self.dims["nsp"])
        _check(lib().ccgpu_get_fock_diag(self.h, _dp(x)))
        return x[0::2] + 1j * x[1::2]

    # reference routine mirrors
    def setup_t_amplitudes(self):
        """MBPT2 start; returns the MBPT2 correlation energy."""
        x = np.zeros(2)
        _check(lib().ccgpu_mbpt2_start(self.h, _dp(x)))
        return complex(x[0], x[1])

    def t2_intermediate(self, precision=PREC_FP64):
        _check(lib().ccgpu_residual(self.h, precision))

    def diis_t1t2(self, count, nstep, subspace=6, diis_step=6):
        c, n = C.c_int(count), C.c_int(nstep)
        _check(lib().ccgpu_update(self.h, C.byref(c), C.byref(n), subspace, diis_step))
        return c.value, n.value

    def ccd_energy_save(self):
        x = np.zeros(2)
        _check(lib().ccgpu_energy(self.h, _dp(x)))
        return complex(x[0], x[1])

    def ccd_iter(self, **kw):
        o = default_opts(**kw)
        r = SolveResult()
        _check(lib().ccgpu_solve(self.h, C.byref(o), C.byref(r)))
        return r.as_dict()

    def solve_energies(self):
        """Correlation energies of the last solve: the start, then one per
        iteration."""
        cnt = C.c_int()
        _check(lib().ccgpu_solve_energies(self.h, None, 0, C.byref(cnt)))
        out = np.zeros(2 * cnt.value)
        _check(lib().ccgpu_solve_energies(self.h, _dp(out), cnt.value, C.byref(cnt)))
        return out[0::2] + 1j * out[1::2]

    def pert_triples2(self):
        x = np.zeros(2)
        _check(lib().ccgpu_triples(self.h, _dp(x)))
        return complex(x[0], x[1])

    def triples_parts(self):
        """(et_4, et_5) of the last pert_triples2 call."""
        a, b = np.zeros(2), np.zeros(2)
        _check(lib().ccgpu_triples_parts(self.h, _dp(a), _dp(b)))
        return complex(a[0], a[1]), complex(b[0], b[1])

    # CCDT-1 (the CCDT1 branch of ccd_iter, ccm_iter.f90:26-92)
    def setup_t3(self):
        x = np.zeros(2)
        _check(lib().ccgpu_setup_t3(self.h, _dp(x)))
        return complex(x[0], x[1])

    def t3_intermediate(self):
        _check(lib().ccgpu_t3_intermediate(self.h))

    def t3_energy(self):
        x = np.zeros(2)
        _check(lib().ccgpu_t3_energy(self.h, _dp(x)))
        return complex(x[0], x[1])

    def triples_stats(self):
        """(elements, seconds) of the last pert_triples2 call."""
        n, t = C.c_longlong(), C.c_double()
        _check(lib().ccgpu_triples_stats(self.h, C.byref(n), C.byref(t)))
        return n.value, t.value

    # amplitudes
    def get_t2(self, a, b, i, j):
        a, b, i, j = (np.ascontiguousarray(x, np.int32) for x in (a, b, i, j))
        out = np.zeros(2 * len(a))
        _check(lib().ccgpu_get_t2(self.h, len(a), _ip(a), _ip(b), _ip(i), _ip(j), _dp(out)))
        return out[0::2] + 1j * out[1::2]

    def set_t2(self, a, b, i, j, t):
        a, b, i, j = (np.ascontiguousarray(x, np.int32) for x in (a, b, i, j))
        tt = np.ascontiguousarray(np.asarray(t, np.complex128))
        _check(lib().ccgpu_set_t2(self.h, len(a), _ip(a), _ip(b), _ip(i), _ip(j), tt.ctypes.data_as(_DP)))

    def t2_unique(self):
        out = np.zeros(self.dims["t2_unique"], np.complex128)
        _check(lib().ccgpu_get_t2_unique(self.h, out.ctypes.data_as(_DP)))
        return out

    def set_t2_unique(self, t):
        tt = np.ascontiguousarray(np.asarray(t, np.complex128))
        assert tt.size == self.dims["t2_unique"]
        _check(lib().ccgpu_set_t2_unique(self.h, tt.ctypes.data_as(_DP)))

    DEBUG = {"R": 0, "dh": 1, "I7off": 2, "dp": 3, "I6off": 4, "X": 5, "Rt": 6, "Tt": 7, "I4": 8,
             "T": 9, "R0": 10, "fdiag": 11, "fpart": 12, "W0": 13, "Vt": 14}

    def debug_fetch(self, what, n):
        out = np.zeros(n, np.complex128)
        _check(lib().ccgpu_debug_fetch(self.h, self.DEBUG[what], out.ctypes.data_as(_DP), n))
        return out

    def t2_labels(self):
        out = np.zeros((self.dims["t2_unique"], 4), np.int32)
        _check(lib().ccgpu_get_t2_labels(self.h, _ip(out)))
        return out


def hbm_read_probe(nbytes=1 << 31, reps=10):
    """Best streaming-read rate (GB/s) of this GPU's HBM over nbytes."""
    x = np.zeros(1)
    _check(lib().ccgpu_hbm_read_probe(int(nbytes), int(reps), _dp(x)))
    return float(x[0])


def batch_solve(solvers, **kw):
    o = default_opts(**kw)
    hs = np.array([s.h for s in solvers], np.int32)
    res = (SolveResult * len(solvers))()
    _check(lib().ccgpu_batch_solve(len(solvers), _ip(hs), C.byref(o), res))
    return [r.as_dict() for r in res]


def energy_noalloc(n, t_ptr, r0_ptr, out_ptr, stream=None):
    """ccgpu_energy_noalloc on device pointers (ints, e.g. torch data_ptr());
    stream-ordered, no synchronization, no allocation."""
    _check(lib().ccgpu_energy_noalloc(int(n), C.c_void_p(t_ptr), C.c_void_p(r0_ptr), C.c_void_p(out_ptr),
                                      C.c_void_p(stream or 0)))


def kernel_stats():
    names = C.create_string_buffer(4096)
    avg = np.zeros(64)
    cnt = np.zeros(64, np.int64)
    k = lib().ccgpu_kernel_stats(names, 4096, _dp(avg), cnt.ctypes.data_as(C.POINTER(C.c_longlong)), 64)
    nm = names.value.decode().split(";") if k else []
    return {nm[x]: (float(avg[x]), int(cnt[x])) for x in range(k)}


def autotune_report():
    buf = C.create_string_buffer(65536)
    lib().ccgpu_autotune_report(buf, 65536)
    return buf.value.decode()


MLP_NPARAM = 9473


class WarmStart:
    """The learned start (ccgpu_mlp_*): trained on converged solves, used in
    place of setup_t_amplitudes."""

    def __init__(self, seed=1):
        h = C.c_int()
        _check(lib().ccgpu_mlp_create(seed, C.byref(h)))
        self.h = h.value

    def __del__(self):
        try:
            if _lib is not None and getattr(self, "h", None):
                _lib.ccgpu_mlp_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass

    @property
    def params(self):
        th = np.zeros(MLP_NPARAM, np.float32)
        _check(lib().ccgpu_mlp_get_params(self.h, th.ctypes.data_as(_FP), MLP_NPARAM))
        return th

    @params.setter
    def params(self, th):
        th = np.ascontiguousarray(th, np.float32)
        _check(lib().ccgpu_mlp_set_params(self.h, th.ctypes.data_as(_FP), len(th)))

    def add_solve(self, solver):
        n = C.c_longlong()
        _check(lib().ccgpu_mlp_add_solve(self.h, solver.h, C.byref(n)))
        return n.value

    def train(self, steps, batch=16384, lr=1e-3, seed=7):
        loss = C.c_double()
        _check(lib().ccgpu_mlp_train(self.h, steps, batch, lr, seed, C.byref(loss)))
        return loss.value

    def warm_start(self, solver):
        x = np.zeros(2)
        _check(lib().ccgpu_mlp_warm_start(self.h, solver.h, _dp(x)))
        return complex(x[0], x[1])

    def warm_start_batch(self, solvers):
        hs = np.array([s.h for s in solvers], np.int32)
        _check(lib().ccgpu_mlp_warm_start_batch(self.h, _ip(hs), len(hs)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(x), np.float32)
        _check(lib().ccgpu_mlp_forward(self.h, len(x), x.ctypes.data_as(_FP), y.ctypes.data_as(_FP)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double()
        _check(lib().ccgpu_mlp_loss_grad(self.h, len(t), x.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                                         C.byref(loss), g.ctypes.data_as(_FP)))
        return loss.value, g

    def features(self, solver, with_target=True):
        n = solver.dims["t2_unique"]
        f = np.zeros((n, 16), np.float32)
        t = np.zeros(n, np.float32) if with_target else None
        _check(lib().ccgpu_mlp_features(self.h, solver.h, f.ctypes.data_as(_FP),
                                        t.ctypes.data_as(_FP) if with_target else None))
        return (f, t) if with_target else f
