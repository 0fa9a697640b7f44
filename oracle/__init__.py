"""ctypes front-end of the CPU oracle (``liborcpf.so``).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C FP64 restatement of the
reference's CPU Newton path (src/powerflow/newtonpf.h:365-523 and the
functions it calls; the full list is in ``pf_oracle.h``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, and only as the checker.  The product package ``mapol_amd`` never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
REF_SO = os.path.join(_HERE, "_ref", "libpfref.so")
_SO = {"own": "liborcpf.so", "x": "liborcpf_x.so"}

PQ, PV, SLACK = 1, 2, 3
NR, FDXB, FDBX = 0, 1, 2
CONVERGED, MAX_ITER, SINGULAR, NONFINITE = 0, 1, 2, 3

_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pu8 = C.POINTER(C.c_ubyte)


class OrcCase(C.Structure):
    _fields_ = [("nbus", C.c_int), ("nbranch", C.c_int), ("ngen", C.c_int),
                ("bus_type", _pi), ("Pd", _pd), ("Qd", _pd), ("Gs", _pd),
                ("Bs", _pd), ("Vm", _pd), ("Va", _pd),
                ("br_from", _pi), ("br_to", _pi), ("R", _pd), ("X", _pd),
                ("B", _pd), ("tap", _pd), ("shift", _pd),
                ("br_status", _pu8), ("gen_bus", _pi), ("Pg", _pd),
                ("Qg", _pd), ("Qmax", _pd), ("Qmin", _pd), ("VG", _pd),
                ("gen_status", _pu8)]


class OrcInst(C.Structure):
    _fields_ = [("br_on", _pu8), ("gen_on", _pu8), ("vset", _pd),
                ("tap", _pd), ("bsh", _pd), ("V0", _pd)]


class OrcOpts(C.Structure):
    _fields_ = [("alg", C.c_int), ("tol", C.c_double), ("maxit_nr", C.c_int),
                ("maxit_fd", C.c_int), ("enforce_qlim", C.c_int),
                ("max_qlim_rounds", C.c_int), ("qlim_tol", C.c_double),
                ("pivot_thresh", C.c_double)]


class OrcResult(C.Structure):
    _fields_ = [("V", _pd), ("bus_type", _pi), ("iterations", C.c_int),
                ("verdict", C.c_int), ("qlim_rounds", C.c_int),
                ("converged", C.c_int), ("norm", C.c_double),
                ("loss", C.c_double)]


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liborcpf.so")
    if force or not all(os.path.exists(os.path.join(_HERE, f)) for f in _SO.values()):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)
    return so


def have_ref() -> bool:
    """True when the Eigen-linked build (oracle/_ref) is present."""
    return os.path.exists(REF_SO)


def lib(which: str = "own"):
    """``own``: the oracle with its own sparse LU; ``ref``: the build whose
    Newton steps go through the reference's Eigen SparseLU; ``x``: the own
    oracle compiled in 80-bit long double (FP64 arrays at the interface)."""
    if which not in _LIBS:
        if which == "ref":
            so = REF_SO
            if not os.path.exists(so):
                raise FileNotFoundError(so)
        else:
            so = os.path.join(_HERE, _SO[which])
            if not os.path.exists(so):
                build()
        L = C.CDLL(so)
        L.orc_solve.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                C.POINTER(OrcOpts), C.POINTER(OrcResult)]
        L.orc_make_ybus.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                    _pi, _pi, _pd, C.c_int]
        L.orc_make_bmat.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                    C.c_int, C.c_int, _pi, _pi, _pi, _pd,
                                    C.c_int]
        L.orc_mismatch.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                   _pd, _pd, _pd]
        L.orc_jacobian.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                   _pd, _pi, _pi, _pd, C.c_int]
        L.orc_initial_voltage.argtypes = [C.POINTER(OrcCase),
                                          C.POINTER(OrcInst), _pd]
        L.orc_objective.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                    _pd, _pd, C.c_double, _pd, _pd]
        L.orc_objective.restype = C.c_double
        L.orc_sparse_solve.argtypes = [C.c_int, _pi, _pi, _pd, _pd, _pd,
                                       C.c_double]
        L.orc_min_degree.argtypes = [C.c_int, _pi, _pi, _pi]
        L.orc_solve_objective.argtypes = [C.POINTER(OrcCase), C.POINTER(OrcInst),
                                          C.POINTER(OrcOpts), _pd, C.c_double,
                                          C.POINTER(OrcResult), _pd, _pd]
        L.orc_precision_bits.restype = C.c_int
        _LIBS[which] = L
    return _LIBS[which]


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Oracle:
    """Holds one case in oracle layout; solves instances one at a time."""

    def __init__(self, case, which: str = "own"):
        self.case = case
        self.which = which
        c = case
        self._keep = {}

        def arr(name, v, dt):
            a = np.ascontiguousarray(v, dtype=dt)
            self._keep[name] = a
            return a

        s = OrcCase()
        s.nbus, s.nbranch, s.ngen = c.nbus, c.nbranch, c.ngen
        s.bus_type = _p(arr("bt", c.bus_type, np.int32), _pi)
        for f in ("Pd", "Qd", "Gs", "Bs", "Vm", "Va"):
            setattr(s, f, _p(arr(f, getattr(c, f), np.float64), _pd))
        s.br_from = _p(arr("f", c.br_from, np.int32), _pi)
        s.br_to = _p(arr("t", c.br_to, np.int32), _pi)
        for f in ("R", "X", "B", "tap", "shift"):
            setattr(s, f, _p(arr(f, getattr(c, f), np.float64), _pd))
        s.br_status = _p(arr("bs", c.br_status, np.uint8), _pu8)
        s.gen_bus = _p(arr("gb", c.gen_bus, np.int32), _pi)
        for f in ("Pg", "Qg", "Qmax", "Qmin", "VG"):
            setattr(s, f, _p(arr("g" + f, getattr(c, f), np.float64), _pd))
        s.gen_status = _p(arr("gs", c.gen_status, np.uint8), _pu8)
        self.s = s

    def _inst(self, br_on=None, gen_on=None, vset=None, tap=None, bsh=None,
              V0=None):
        keep = []

        def ptr(v, dt, t):
            if v is None:
                return None
            a = np.ascontiguousarray(v, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(t)

        i = OrcInst(ptr(br_on, np.uint8, _pu8), ptr(gen_on, np.uint8, _pu8),
                    ptr(vset, np.float64, _pd), ptr(tap, np.float64, _pd),
                    ptr(bsh, np.float64, _pd),
                    ptr(None if V0 is None else
                        np.stack([np.real(V0), np.imag(V0)], -1).ravel(),
                        np.float64, _pd))
        return i, keep

    @staticmethod
    def opts(alg=NR, tol=1e-8, maxit_nr=10, maxit_fd=30, enforce_qlim=False,
             max_qlim_rounds=10, qlim_tol=1e-6, pivot_thresh=1.0):
        return OrcOpts(alg, tol, maxit_nr, maxit_fd, int(enforce_qlim),
                       max_qlim_rounds, qlim_tol, pivot_thresh)

    def solve(self, opts=None, **inst):
        o = opts or self.opts()
        n = self.case.nbus
        V = np.zeros(2 * n)
        bt = np.zeros(n, dtype=np.int32)
        r = OrcResult(_p(V, _pd), _p(bt, _pi), 0, 0, 0, 0, 0.0, 0.0)
        i, keep = self._inst(**inst)
        lib(self.which).orc_solve(C.byref(self.s), C.byref(i), C.byref(o), C.byref(r))
        return dict(V=V[0::2] + 1j * V[1::2], bus_type=bt,
                    iterations=r.iterations, verdict=r.verdict,
                    qlim_rounds=r.qlim_rounds, converged=bool(r.converged),
                    norm=r.norm, loss=r.loss)

    def solve_objective(self, opts=None, rate=None, w=0.0, **inst):
        """solve, then the objective at the solution in the working precision
        (returned unrounded as the pair (hi, lo), objective = hi + lo)"""
        o = opts or self.opts()
        n = self.case.nbus
        V = np.zeros(2 * n)
        bt = np.zeros(n, dtype=np.int32)
        r = OrcResult(_p(V, _pd), _p(bt, _pi), 0, 0, 0, 0, 0.0, 0.0)
        hi, lo = np.zeros(1), np.zeros(1)
        rt = None if rate is None else np.ascontiguousarray(rate, np.float64)
        i, keep = self._inst(**inst)
        lib(self.which).orc_solve_objective(C.byref(self.s), C.byref(i), C.byref(o),
                                            _p(rt, _pd), w, C.byref(r),
                                            _p(hi, _pd), _p(lo, _pd))
        return dict(V=V[0::2] + 1j * V[1::2], iterations=r.iterations,
                    verdict=r.verdict, converged=bool(r.converged), norm=r.norm,
                    f_hi=float(hi[0]), f_lo=float(lo[0]))

    def ybus(self, **inst):
        n = self.case.nbus
        cap = n + 2 * self.case.nbranch
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(cap, np.int32)
        v = np.zeros(2 * cap)
        i, keep = self._inst(**inst)
        nnz = lib(self.which).orc_make_ybus(C.byref(self.s), C.byref(i), _p(rp, _pi),
                                  _p(ci, _pi), _p(v, _pd), cap)
        assert nnz >= 0
        import scipy.sparse as sp
        return sp.csr_matrix((v[0:2 * nnz:2] + 1j * v[1:2 * nnz:2],
                              ci[:nnz], rp), shape=(n, n))

    def bmat(self, alg, which, **inst):
        """B' (which 1, over pv then pq) or B'' (which 2, over pq) of the
        fast-decoupled variant ``alg`` as a scipy CSR matrix"""
        import scipy.sparse as sp
        n = self.case.nbus
        cap = n + 2 * self.case.nbranch
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        rows = np.zeros(1, np.int32)
        i, keep = self._inst(**inst)
        nnz = lib(self.which).orc_make_bmat(C.byref(self.s), C.byref(i), alg,
                                            which, _p(rows, _pi), _p(rp, _pi),
                                            _p(ci, _pi), _p(v, _pd), cap)
        assert nnz >= 0
        rows = int(rows[0])
        return sp.csr_matrix((v[:nnz], ci[:nnz], rp[:rows + 1]),
                             shape=(rows, rows))

    def initial_voltage(self, **inst):
        n = self.case.nbus
        V = np.zeros(2 * n)
        i, keep = self._inst(**inst)
        lib(self.which).orc_initial_voltage(C.byref(self.s), C.byref(i), _p(V, _pd))
        return V[0::2] + 1j * V[1::2]

    def mismatch(self, V, **inst):
        n = self.case.nbus
        Vd = np.ascontiguousarray(np.stack([V.real, V.imag], -1).ravel())
        F = np.zeros(2 * n)
        nrm = np.zeros(1)
        i, keep = self._inst(**inst)
        nJ = lib(self.which).orc_mismatch(C.byref(self.s), C.byref(i), _p(Vd, _pd),
                                _p(F, _pd), _p(nrm, _pd))
        return F[:nJ], float(nrm[0])

    def jacobian(self, V, **inst):
        import scipy.sparse as sp
        n = self.case.nbus
        Vd = np.ascontiguousarray(np.stack([V.real, V.imag], -1).ravel())
        cap = 4 * (n + 2 * self.case.nbranch)
        rp = np.zeros(2 * n + 1, np.int32)
        ci = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        i, keep = self._inst(**inst)
        nnz = lib(self.which).orc_jacobian(C.byref(self.s), C.byref(i), _p(Vd, _pd),
                                 _p(rp, _pi), _p(ci, _pi), _p(v, _pd), cap)
        assert nnz >= 0
        F, _ = self.mismatch(V, **inst)
        nJ = len(F)
        return sp.csr_matrix((v[:nnz], ci[:nnz], rp[:nJ + 1]),
                             shape=(nJ, nJ))

    def objective(self, V, rate=None, w=0.0, **inst):
        m = self.case.nbranch
        Vd = np.ascontiguousarray(np.stack([V.real, V.imag], -1).ravel())
        Sf = np.zeros(2 * m)
        St = np.zeros(2 * m)
        rt = None if rate is None else np.ascontiguousarray(rate, np.float64)
        i, keep = self._inst(**inst)
        f = lib(self.which).orc_objective(C.byref(self.s), C.byref(i), _p(Vd, _pd),
                                _p(rt, _pd), w, _p(Sf, _pd), _p(St, _pd))
        return f, Sf[0::2] + 1j * Sf[1::2], St[0::2] + 1j * St[1::2]


def sparse_solve(A, b, pivot_thresh=1.0):
    """Solve with the oracle's sparse LU (A: scipy CSR, float64)."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, np.int32)
    ci = np.ascontiguousarray(A.indices, np.int32)
    v = np.ascontiguousarray(A.data, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    st = lib().orc_sparse_solve(n, _p(rp, _pi), _p(ci, _pi), _p(v, _pd),
                                _p(bb, _pd), _p(x, _pd), pivot_thresh)
    return x, st


def min_degree(A):
    A = A.tocsr()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, np.int32)
    ci = np.ascontiguousarray(A.indices, np.int32)
    p = np.zeros(n, np.int32)
    lib().orc_min_degree(n, _p(rp, _pi), _p(ci, _pi), _p(p, _pi))
    return p
