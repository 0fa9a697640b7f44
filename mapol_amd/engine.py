"""ctypes binding of ``libbatchpf.so`` (declared in include/batchpf.h).

The product path: every call goes through the C-ABI into the HIP engine.
There is no CPU fallback; a missing library or device is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbatchpf.so")

BPF_NR, BPF_FDXB, BPF_FDBX = 0, 1, 2
BPF_FP64, BPF_MIXED = 0, 1
CONVERGED, MAX_ITER, SINGULAR, NONFINITE = 0, 1, 2, 3
CTL_AVR, CTL_OLTC, CTL_SHC = 1, 2, 3

_pi = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)
_pu8 = C.POINTER(C.c_uint8)

# every entry point of include/batchpf.h
EXPORTS = [
    "bpf_default_options", "bpf_create", "bpf_set_instances", "bpf_solve",
    "bpf_get_results", "bpf_get_branch_flows", "bpf_get_branch_loading",
    "bpf_get_branch_loading_noalloc", "bpf_destroy",
    "bpf_last_error", "bpf_get_info", "bpf_make_ybus", "bpf_get_ybus",
    "bpf_stage_bmat", "bpf_stage_fd_halfstep", "bpf_stage_mismatch",
    "bpf_stage_jacobian", "bpf_stage_newton_step", "bpf_stage_tail_inverse",
    "bpf_adjoint",
    "bpf_set_timing", "bpf_get_kernel_time", "bpf_set_launch",
    "bpf_set_schedule", "bpf_set_factor_split", "bpf_set_fused_forward",
    "bpf_analyse", "bpf_analyse_sweeps", "bpf_analyse_fill", "bpf_analyse_tail",
    "bpf_analyse_slots", "bpf_analyse_slots_transposed",
    "bpf_load_matpower",
    "bpf_case_error", "bpf_free_case",
]


class BpfCase(C.Structure):
    _fields_ = [("nbus", C.c_int32), ("nbranch", C.c_int32),
                ("ngen", C.c_int32), ("baseMVA", C.c_double),
                ("bus_type", _pi), ("Pd", _pd), ("Qd", _pd), ("Gs", _pd),
                ("Bs", _pd), ("Vm", _pd), ("Va", _pd), ("br_from", _pi),
                ("br_to", _pi), ("R", _pd), ("X", _pd), ("B", _pd),
                ("tap", _pd), ("shift", _pd), ("br_status", _pu8),
                ("gen_bus", _pi), ("Pg", _pd), ("Qg", _pd), ("Qmax", _pd),
                ("Qmin", _pd), ("VG", _pd), ("gen_status", _pu8)]


class BpfOptions(C.Structure):
    _fields_ = [("alg", C.c_int32), ("precision", C.c_int32),
                ("refine_steps", C.c_int32), ("maxit_nr", C.c_int32),
                ("maxit_fd", C.c_int32), ("enforce_qlim", C.c_int32),
                ("max_qlim_rounds", C.c_int32), ("use_graph", C.c_int32),
                ("tol", C.c_double), ("qlim_tol", C.c_double),
                ("tail_limit", C.c_int32), ("chunk_products", C.c_int32),
                ("slot_chunk", C.c_int32), ("ordering", C.c_int32),
                ("bus_numbering", C.c_int32)]


class BpfCtl(C.Structure):
    _fields_ = [("kind", C.c_int32), ("id", C.c_int32)]


class BpfInfo(C.Structure):
    _fields_ = [("nbus", C.c_int32), ("nbranch", C.c_int32),
                ("nblocks", C.c_int32), ("nnz_ybus", C.c_int32),
                ("nnz_jac_blocks", C.c_int32), ("nnz_lu_blocks", C.c_int32),
                ("lu_levels", C.c_int32), ("fw_levels", C.c_int32),
                ("bw_levels", C.c_int32), ("lu_pairs", C.c_int64),
                ("instances", C.c_int32), ("lds_solve", C.c_int32),
                ("crout_tasks", C.c_int32), ("chunks", C.c_int32),
                ("graph_used", C.c_int32), ("tail_level", C.c_int32),
                ("tail_n", C.c_int32), ("tail_entries", C.c_int32),
                ("tail_pairs", C.c_int32), ("tail_row_entries", C.c_int32),
                ("sweep_fw_chunks", C.c_int32), ("sweep_bw_chunks", C.c_int32),
                ("sweep_rows", C.c_int32), ("sweep_entries", C.c_int32),
                ("fill_level0_chunks", C.c_int32), ("tail_stage", C.c_int32),
                ("slot_fw_chunks", C.c_int32), ("slot_bw_chunks", C.c_int32),
                ("slots", C.c_int32), ("graph_fallback", C.c_int32),
                ("graph_captures", C.c_int32), ("fused_forward", C.c_int32)]


_LIB = None


class EngineError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load the engine library (build it with ``python -m mapol_amd.build``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise EngineError(f"{path} is missing: build it with "
                          "`python -m mapol_amd.build`")
    L = C.CDLL(path)
    V = C.c_void_p
    L.bpf_default_options.argtypes = [C.POINTER(BpfOptions)]
    L.bpf_default_options.restype = None
    L.bpf_create.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfOptions),
                             C.POINTER(V)]
    L.bpf_set_instances.argtypes = [V, C.c_int32, _pu8, _pu8, C.c_int32,
                                    C.POINTER(BpfCtl), _pd]
    L.bpf_solve.argtypes = [V]
    L.bpf_get_results.argtypes = [V, _pd, _pi, _pi, _pd, _pd, _pi, _pi]
    L.bpf_get_branch_flows.argtypes = [V, _pd, _pd]
    L.bpf_get_branch_loading.argtypes = [V, _pd, _pd, _pi, _pd]
    L.bpf_get_branch_loading_noalloc.argtypes = [V, _pd, _pd, _pi, _pd]
    L.bpf_destroy.argtypes = [V]
    L.bpf_destroy.restype = None
    L.bpf_last_error.argtypes = [V]
    L.bpf_last_error.restype = C.c_char_p
    L.bpf_get_info.argtypes = [V, C.POINTER(BpfInfo)]
    L.bpf_make_ybus.argtypes = [V]
    L.bpf_get_ybus.argtypes = [V, C.c_int32, _pi, _pi, _pd]
    L.bpf_stage_bmat.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, _pi, _pi,
                                 _pd, _pi]
    L.bpf_stage_fd_halfstep.argtypes = [V, _pd, C.c_int32, _pd, _pd]
    L.bpf_stage_mismatch.argtypes = [V, _pd, _pd, _pd]
    L.bpf_stage_jacobian.argtypes = [V, _pd, C.c_int32, C.c_int32, _pi, _pi,
                                     _pd, _pi]
    L.bpf_stage_newton_step.argtypes = [V, _pd, _pd]
    L.bpf_stage_tail_inverse.argtypes = [V, C.c_int32, _pi, _pd]
    L.bpf_adjoint.argtypes = [V, _pd, C.c_double, _pd, _pd, _pd, _pd]
    L.bpf_set_timing.argtypes = [V, C.c_int32]
    L.bpf_get_kernel_time.argtypes = [V, C.c_char_p, _pi, _pd]
    L.bpf_set_launch.argtypes = [V, C.c_int32, C.c_int32, C.c_int32]
    L.bpf_set_schedule.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.bpf_set_factor_split.argtypes = [V, C.c_int32]
    L.bpf_set_fused_forward.argtypes = [V, C.c_int32]
    L.bpf_analyse.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfInfo)] + [_pi] * 21
    L.bpf_analyse_sweeps.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfInfo)] + [_pi] * 3
    L.bpf_analyse_fill.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfInfo)] + [_pi] * 2
    L.bpf_analyse_tail.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfInfo)] + [_pi] * 2
    L.bpf_analyse_slots.argtypes = [C.POINTER(BpfCase), C.POINTER(BpfInfo)] + [_pi] * 2
    L.bpf_analyse_slots_transposed.argtypes = [C.POINTER(BpfCase)] + [_pi] * 3
    L.bpf_load_matpower.argtypes = [C.c_char_p, C.c_int32, C.POINTER(BpfCase),
                                    C.POINTER(_pd), C.POINTER(C.c_void_p)]
    L.bpf_case_error.argtypes = [C.c_void_p]
    L.bpf_case_error.restype = C.c_char_p
    L.bpf_free_case.argtypes = [C.c_void_p]
    L.bpf_free_case.restype = None
    for name in EXPORTS:
        getattr(L, name)
    _LIB = L
    return L


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _case_struct(c, keep):
    def arr(v, dt):
        a = np.ascontiguousarray(v, dtype=dt)
        keep.append(a)
        return a

    s = BpfCase()
    s.nbus, s.nbranch, s.ngen = c.nbus, c.nbranch, c.ngen
    s.baseMVA = float(getattr(c, "baseMVA", 100.0))
    s.bus_type = _ptr(arr(c.bus_type, np.int32), _pi)
    for f in ("Pd", "Qd", "Gs", "Bs", "Vm", "Va", "R", "X", "B", "tap",
              "shift", "Pg", "Qg", "Qmax", "Qmin", "VG"):
        setattr(s, f, _ptr(arr(getattr(c, f), np.float64), _pd))
    s.br_from = _ptr(arr(c.br_from, np.int32), _pi)
    s.br_to = _ptr(arr(c.br_to, np.int32), _pi)
    s.br_status = _ptr(arr(c.br_status, np.uint8), _pu8)
    s.gen_bus = _ptr(arr(c.gen_bus, np.int32), _pi)
    s.gen_status = _ptr(arr(c.gen_status, np.uint8), _pu8)
    return s


BPF_MP_GEN_STATUS = 1


def load_matpower(path: str, honour_gen_status: bool = False):
    """Read a MATPOWER version-2 case file through ``bpf_load_matpower``
    (host only; the reference's loadMatpowerFormat conventions).  Returns
    ``(case, rate_a)``: a :class:`mapol_amd.synth.Case` with copies of the
    arrays and the branch ratings (rateA / baseMVA)."""
    from .synth import Case

    L = load_library()
    s = BpfCase()
    rate = _pd()
    owner = C.c_void_p()
    st = L.bpf_load_matpower(os.fsencode(path),
                             BPF_MP_GEN_STATUS if honour_gen_status else 0,
                             C.byref(s), C.byref(rate), C.byref(owner))
    try:
        if st != 0:
            msg = L.bpf_case_error(owner).decode() if owner.value else ""
            raise EngineError(f"bpf_load_matpower failed ({st}): {msg}")
        n, m, g = s.nbus, s.nbranch, s.ngen

        def take(ptr, k, dt):
            return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt, copy=True) \
                if k else np.zeros(0, dt)

        case = Case(
            bus_type=take(s.bus_type, n, np.int32),
            Pd=take(s.Pd, n, np.float64), Qd=take(s.Qd, n, np.float64),
            Gs=take(s.Gs, n, np.float64), Bs=take(s.Bs, n, np.float64),
            Vm=take(s.Vm, n, np.float64), Va=take(s.Va, n, np.float64),
            br_from=take(s.br_from, m, np.int32), br_to=take(s.br_to, m, np.int32),
            R=take(s.R, m, np.float64), X=take(s.X, m, np.float64),
            B=take(s.B, m, np.float64), tap=take(s.tap, m, np.float64),
            shift=take(s.shift, m, np.float64),
            br_status=take(s.br_status, m, np.uint8),
            gen_bus=take(s.gen_bus, g, np.int32), Pg=take(s.Pg, g, np.float64),
            Qg=take(s.Qg, g, np.float64), Qmax=take(s.Qmax, g, np.float64),
            Qmin=take(s.Qmin, g, np.float64), VG=take(s.VG, g, np.float64),
            gen_status=take(s.gen_status, g, np.uint8), baseMVA=s.baseMVA,
            name=os.path.splitext(os.path.basename(path))[0],
            rate=take(rate, m, np.float64))
        return case, case.rate
    finally:
        if owner.value:
            L.bpf_free_case(owner)


def analyse(case):
    """Host-only symbolic analysis (``bpf_analyse``): block order, L+U block
    pattern and level schedules, as numpy arrays (no GPU needed)."""
    L = load_library()
    keep = []
    s = _case_struct(case, keep)
    info = BpfInfo()
    st = L.bpf_analyse(C.byref(s), C.byref(info), *([None] * 21))
    if st != 0:
        raise EngineError(f"bpf_analyse failed ({st})")
    i = {f: getattr(info, f) for f, _ in BpfInfo._fields_}
    nb, nl = i["nblocks"], i["lu_levels"]
    out = dict(
        bus_of_blk=np.zeros(nb, np.int32), lu_rowptr=np.zeros(nb + 1, np.int32),
        lu_colind=np.zeros(i["nnz_lu_blocks"], np.int32),
        ch_task=np.zeros(i["chunks"] + 1, np.int32),
        ch_pair=np.zeros(i["chunks"] + 1, np.int32),
        tk_a=np.zeros((i["crout_tasks"], 4), np.int32),
        tk_b=np.zeros((i["crout_tasks"], 4), np.int32),
        pairs=np.zeros((i["lu_pairs"], 2), np.int32),
        fw_ptr=np.zeros(i["fw_levels"] + 1, np.int32),
        fw_rows=np.zeros(nb - i["tail_n"], np.int32),
        bw_ptr=np.zeros(i["bw_levels"] + 1, np.int32),
        bw_rows=np.zeros(nb - i["tail_n"], np.int32),
        tl_dense=np.zeros(i["tail_entries"], np.int32),
        tl_pos=np.zeros(i["tail_entries"], np.int32),
        tl_ptr=np.zeros(i["tail_entries"] + 1, np.int32),
        tl_pairL=np.zeros(i["tail_pairs"], np.int32),
        tl_pairU=np.zeros(i["tail_pairs"], np.int32),
        tail_piv=np.zeros(i["tail_n"], np.int32),
        ts_ptr=np.zeros(i["tail_n"] + 1, np.int32),
        ts_pos=np.zeros(i["tail_row_entries"], np.int32),
        ts_col=np.zeros(i["tail_row_entries"], np.int32))
    order = ["bus_of_blk", "lu_rowptr", "lu_colind", "ch_task", "ch_pair",
             "tk_a", "tk_b", "pairs", "fw_ptr", "fw_rows",
             "bw_ptr", "bw_rows", "tl_dense", "tl_pos", "tl_ptr", "tl_pairL",
             "tl_pairU", "tail_piv", "ts_ptr", "ts_pos", "ts_col"]
    st = L.bpf_analyse(C.byref(s), C.byref(info),
                       *[_ptr(out[k], _pi) for k in order])
    if st != 0:
        raise EngineError(f"bpf_analyse failed ({st})")
    out["info"] = i
    nch = i["sweep_fw_chunks"] + i["sweep_bw_chunks"]
    sw = dict(sw_chunk=np.zeros((nch, 4), np.int32),
              sw_row=np.zeros((i["sweep_rows"], 4), np.int32),
              sw_ent=np.zeros((i["sweep_entries"], 2), np.int32))
    st = L.bpf_analyse_sweeps(C.byref(s), C.byref(info), _ptr(sw["sw_chunk"], _pi),
                              _ptr(sw["sw_row"], _pi), _ptr(sw["sw_ent"], _pi))
    if st != 0:
        raise EngineError(f"bpf_analyse_sweeps failed ({st})")
    out.update(sw)
    fl = dict(jac_lu=np.zeros(i["nnz_jac_blocks"], np.int32),
              jac_l0=np.zeros(i["nnz_jac_blocks"], np.int32))
    st = L.bpf_analyse_fill(C.byref(s), C.byref(info), _ptr(fl["jac_lu"], _pi),
                            _ptr(fl["jac_l0"], _pi))
    if st != 0:
        raise EngineError(f"bpf_analyse_fill failed ({st})")
    out.update(fl)
    tl = dict(tl_stage=np.zeros(i["tail_stage"], np.int32),
              tl_pairS=np.zeros(i["tail_pairs"] if i["tail_stage"] else 0, np.int32))
    st = L.bpf_analyse_tail(C.byref(s), C.byref(info), _ptr(tl["tl_stage"], _pi),
                            _ptr(tl["tl_pairS"], _pi))
    if st != 0:
        raise EngineError(f"bpf_analyse_tail failed ({st})")
    out.update(tl)
    sl = dict(sl_chunk=np.zeros((i["slot_fw_chunks"] + i["slot_bw_chunks"], 4), np.int32),
              sl_slot=np.zeros((i["slots"], 4), np.int32))
    st = L.bpf_analyse_slots(C.byref(s), C.byref(info), _ptr(sl["sl_chunk"], _pi),
                             _ptr(sl["sl_slot"], _pi))
    if st != 0:
        raise EngineError(f"bpf_analyse_slots failed ({st})")
    out.update(sl)
    cnt = np.zeros(3, np.int32)
    st = L.bpf_analyse_slots_transposed(C.byref(s), _ptr(cnt, _pi), None, None)
    if st != 0:
        raise EngineError(f"bpf_analyse_slots_transposed failed ({st})")
    slt = dict(sl_chunk_t=np.zeros((cnt[0] + cnt[1], 4), np.int32),
               sl_slot_t=np.zeros((cnt[2], 4), np.int32))
    st = L.bpf_analyse_slots_transposed(C.byref(s), _ptr(cnt, _pi),
                                        _ptr(slt["sl_chunk_t"], _pi),
                                        _ptr(slt["sl_slot_t"], _pi))
    if st != 0:
        raise EngineError(f"bpf_analyse_slots_transposed failed ({st})")
    out.update(slt)
    out["slot_t_fw_chunks"] = int(cnt[0])
    return out


class Engine:
    """One batched power-flow engine over one case (``synth.Case``)."""

    def __init__(self, case, alg: int = BPF_NR, precision: int = BPF_FP64,
                 refine_steps: int = 1, tol: float = 1e-8, maxit_nr: int = 10,
                 maxit_fd: int = 30, enforce_qlim: bool = False,
                 max_qlim_rounds: int = 10, qlim_tol: float = 1e-6,
                 use_graph: bool = True, tail_limit: int = 0,
                 chunk_products: int = 0, slot_chunk: int = 0,
                 ordering: int = 0, bus_numbering: int = 1):
        self.L = load_library()
        self.case = case
        self._keep = []
        c = case

        def arr(v, dt):
            a = np.ascontiguousarray(v, dtype=dt)
            self._keep.append(a)
            return a

        s = BpfCase()
        s.nbus, s.nbranch, s.ngen = c.nbus, c.nbranch, c.ngen
        s.baseMVA = float(getattr(c, "baseMVA", 100.0))
        s.bus_type = _ptr(arr(c.bus_type, np.int32), _pi)
        for f in ("Pd", "Qd", "Gs", "Bs", "Vm", "Va", "R", "X", "B", "tap",
                  "shift", "Pg", "Qg", "Qmax", "Qmin", "VG"):
            setattr(s, f, _ptr(arr(getattr(c, f), np.float64), _pd))
        s.br_from = _ptr(arr(c.br_from, np.int32), _pi)
        s.br_to = _ptr(arr(c.br_to, np.int32), _pi)
        s.br_status = _ptr(arr(c.br_status, np.uint8), _pu8)
        s.gen_bus = _ptr(arr(c.gen_bus, np.int32), _pi)
        s.gen_status = _ptr(arr(c.gen_status, np.uint8), _pu8)
        o = BpfOptions()
        self.L.bpf_default_options(C.byref(o))
        o.alg, o.precision, o.refine_steps = alg, precision, refine_steps
        o.maxit_nr, o.maxit_fd, o.tol = maxit_nr, maxit_fd, tol
        o.enforce_qlim, o.max_qlim_rounds = int(enforce_qlim), max_qlim_rounds
        o.qlim_tol, o.use_graph = qlim_tol, int(use_graph)
        o.tail_limit, o.chunk_products = int(tail_limit), int(chunk_products)
        o.slot_chunk, o.ordering = int(slot_chunk), int(ordering)
        o.bus_numbering = int(bus_numbering)
        self.options = o
        h = C.c_void_p()
        st = self.L.bpf_create(C.byref(s), C.byref(o), C.byref(h))
        self.h = h
        if st != 0:
            msg = self.L.bpf_last_error(h).decode() if h else "bpf_create"
            self.close()
            raise EngineError(f"bpf_create failed ({st}): {msg}")
        self.T = 0

    # -- helpers -----------------------------------------------------------
    def _check(self, st, what):
        if st != 0:
            raise EngineError(f"{what} failed ({st}): "
                              f"{self.L.bpf_last_error(self.h).decode()}")

    def last_error(self) -> str:
        return self.L.bpf_last_error(self.h).decode()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.bpf_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- API ---------------------------------------------------------------
    def set_instances(self, br_on=None, gen_on=None, T=None, ctl=None,
                      ctl_val=None):
        if br_on is not None:
            br_on = np.ascontiguousarray(br_on, np.uint8)
            T = br_on.shape[0]
        if gen_on is not None:
            gen_on = np.ascontiguousarray(gen_on, np.uint8)
            T = gen_on.shape[0]
        if T is None:
            raise ValueError("T unknown")
        nctl = 0
        cptr = vptr = None
        if ctl is not None:
            nctl = len(ctl)
            carr = (BpfCtl * nctl)(*[BpfCtl(int(k), int(i)) for k, i in ctl])
            self._ctl = carr
            cptr = carr
            vals = np.ascontiguousarray(ctl_val, np.float64).reshape(T, nctl)
            self._ctlv = vals
            vptr = _ptr(vals, _pd)
        self._check(self.L.bpf_set_instances(
            self.h, int(T), _ptr(br_on, _pu8), _ptr(gen_on, _pu8), nctl,
            cptr, vptr), "bpf_set_instances")
        self.T = int(T)

    def solve(self):
        self._check(self.L.bpf_solve(self.h), "bpf_solve")

    def results(self):
        T, n = self.T, self.case.nbus
        V = np.zeros((T, n, 2))
        its = np.zeros(T, np.int32)
        ver = np.zeros(T, np.int32)
        nrm = np.zeros(T)
        loss = np.zeros(T)
        rnd = np.zeros(T, np.int32)
        bt = np.zeros((T, n), np.int32)
        self._check(self.L.bpf_get_results(
            self.h, _ptr(V, _pd), _ptr(its, _pi), _ptr(ver, _pi),
            _ptr(nrm, _pd), _ptr(loss, _pd), _ptr(rnd, _pi), _ptr(bt, _pi)),
            "bpf_get_results")
        return dict(V=V[..., 0] + 1j * V[..., 1], iterations=its, verdict=ver,
                    norm=nrm, loss=loss, qlim_rounds=rnd, bus_type=bt)

    def branch_flows(self):
        T, m = self.T, self.case.nbranch
        Sf = np.zeros((T, m, 2))
        St = np.zeros((T, m, 2))
        self._check(self.L.bpf_get_branch_flows(self.h, _ptr(Sf, _pd),
                                                _ptr(St, _pd)),
                    "bpf_get_branch_flows")
        return Sf[..., 0] + 1j * Sf[..., 1], St[..., 0] + 1j * St[..., 1]

    def branch_loading(self, rate):
        """Per-branch loading (% of ``rate``, pu; 0 = unrated), and per
        instance the number of branches above 100 % and the largest loading,
        at the last solution (``bpf_get_branch_loading``)."""
        T, m = self.T, self.case.nbranch
        rt = np.ascontiguousarray(rate, np.float64)
        if rt.shape != (m,):
            raise EngineError(f"rate must have shape ({m},)")
        load = np.zeros((T, m))
        over = np.zeros(T, np.int32)
        mx = np.zeros(T)
        self._check(self.L.bpf_get_branch_loading(self.h, _ptr(rt, _pd),
                                                  _ptr(load, _pd), _ptr(over, _pi),
                                                  _ptr(mx, _pd)),
                    "bpf_get_branch_loading")
        return dict(loading=load, overloads=over, max_loading=mx)

    def branch_loading_noalloc(self, rate, per_branch=True):
        """The same report through ``bpf_get_branch_loading_noalloc``, which
        allocates no device memory (``loading`` is None when ``per_branch``
        is False)."""
        T, m = self.T, self.case.nbranch
        rt = np.ascontiguousarray(rate, np.float64)
        if rt.shape != (m,):
            raise EngineError(f"rate must have shape ({m},)")
        load = np.zeros((T, m)) if per_branch else None
        over = np.zeros(T, np.int32)
        mx = np.zeros(T)
        self._check(self.L.bpf_get_branch_loading_noalloc(
            self.h, _ptr(rt, _pd), _ptr(load, _pd), _ptr(over, _pi),
            _ptr(mx, _pd)), "bpf_get_branch_loading_noalloc")
        return dict(loading=load, overloads=over, max_loading=mx)

    def info(self):
        i = BpfInfo()
        self._check(self.L.bpf_get_info(self.h, C.byref(i)), "bpf_get_info")
        return {f: getattr(i, f) for f, _ in BpfInfo._fields_}

    def ybus(self, t):
        import scipy.sparse as sp
        n = self.case.nbus
        inf = self.info()
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(inf["nnz_ybus"], np.int32)
        v = np.zeros(2 * inf["nnz_ybus"])
        self._check(self.L.bpf_get_ybus(self.h, int(t), _ptr(rp, _pi),
                                        _ptr(ci, _pi), _ptr(v, _pd)),
                    "bpf_get_ybus")
        return sp.csr_matrix((v[0::2] + 1j * v[1::2], ci, rp), shape=(n, n))

    def make_ybus(self):
        self._check(self.L.bpf_make_ybus(self.h), "bpf_make_ybus")

    def stage_bmat(self, t, which):
        """B' (which 1, pv then pq) or B'' (which 2, pq) of instance t"""
        import scipy.sparse as sp
        cap = self.info()["nnz_ybus"]
        r = np.zeros(cap, np.int32)
        c = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        nnz = C.c_int32()
        self._check(self.L.bpf_stage_bmat(self.h, int(t), int(which), cap,
                                          _ptr(r, _pi), _ptr(c, _pi), _ptr(v, _pd),
                                          C.byref(nnz)), "bpf_stage_bmat")
        k = nnz.value
        nr = int(max(r[:k].max(), c[:k].max())) + 1 if k else 0
        return sp.coo_matrix((v[:k], (r[:k], c[:k])), shape=(nr, nr)).tocsr()

    def stage_fd_halfstep(self, V, which):
        """one P (which 1) or Q (which 2) half-step of every instance from V;
        returns (new V, inf-norm of the |V|-normalised mismatch at V)"""
        T, n = self.T, self.case.nbus
        Vd = self._vin(V)
        out = np.zeros((T, n, 2))
        nrm = np.zeros(T)
        self._check(self.L.bpf_stage_fd_halfstep(self.h, _ptr(Vd, _pd), int(which),
                                                 _ptr(out, _pd), _ptr(nrm, _pd)),
                    "bpf_stage_fd_halfstep")
        return out[..., 0] + 1j * out[..., 1], nrm

    @staticmethod
    def _vin(V):
        V = np.asarray(V)
        return np.ascontiguousarray(np.stack([V.real, V.imag], -1),
                                    np.float64)

    def stage_mismatch(self, V):
        T, n = self.T, self.case.nbus
        Vd = self._vin(V)
        F = np.zeros((T, 2 * n))
        nrm = np.zeros(T)
        self._check(self.L.bpf_stage_mismatch(self.h, _ptr(Vd, _pd),
                                              _ptr(F, _pd), _ptr(nrm, _pd)),
                    "bpf_stage_mismatch")
        return F, nrm

    def stage_jacobian(self, V, t):
        import scipy.sparse as sp
        Vd = self._vin(V)
        cap = 4 * self.info()["nnz_ybus"]
        r = np.zeros(cap, np.int32)
        c = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        nnz = C.c_int32()
        self._check(self.L.bpf_stage_jacobian(
            self.h, _ptr(Vd, _pd), int(t), cap, _ptr(r, _pi), _ptr(c, _pi),
            _ptr(v, _pd), C.byref(nnz)), "bpf_stage_jacobian")
        k = nnz.value
        nJ = int(max(r[:k].max(), c[:k].max())) + 1 if k else 0
        return sp.coo_matrix((v[:k], (r[:k], c[:k])), shape=(nJ, nJ)).tocsr()

    def stage_newton_step(self, V):
        T, n = self.T, self.case.nbus
        Vd = self._vin(V)
        dx = np.zeros((T, 2 * n))
        self._check(self.L.bpf_stage_newton_step(self.h, _ptr(Vd, _pd),
                                                 _ptr(dx, _pd)),
                    "bpf_stage_newton_step")
        return dx

    def stage_tail_inverse(self, t):
        """(buses of the dense tail's pivots, inverse of its Schur complement
        as a dense (2 tail_n)^2 matrix) of instance t from the last
        factorization"""
        s = self.info()["tail_n"]
        bus = np.zeros(s, np.int32)
        Si = np.zeros((2 * s, 2 * s))
        self._check(self.L.bpf_stage_tail_inverse(self.h, int(t), _ptr(bus, _pi),
                                                  _ptr(Si, _pd)),
                    "bpf_stage_tail_inverse")
        return bus, Si

    def adjoint(self, rate=None, w=0.0):
        T, n, m = self.T, self.case.nbus, self.case.nbranch
        gP = np.zeros((T, n))
        gQ = np.zeros((T, n))
        gx = np.zeros((T, m))
        obj = np.zeros(T)
        r = None if rate is None else np.ascontiguousarray(rate, np.float64)
        self._check(self.L.bpf_adjoint(self.h, _ptr(r, _pd), float(w),
                                       _ptr(gP, _pd), _ptr(gQ, _pd),
                                       _ptr(gx, _pd), _ptr(obj, _pd)),
                    "bpf_adjoint")
        return dict(grad_P=gP, grad_Q=gQ, grad_x=gx, objective=obj)

    def set_timing(self, on: bool):
        self._check(self.L.bpf_set_timing(self.h, int(on)), "bpf_set_timing")

    def kernel_time(self, name: str):
        n = C.c_int32()
        ms = C.c_double()
        self._check(self.L.bpf_get_kernel_time(self.h, name.encode(),
                                               C.byref(n), C.byref(ms)),
                    "bpf_get_kernel_time")
        return n.value, ms.value

    def set_schedule(self, newton_per_graph=0, tail_inverse=0, sweep_form=0,
                     parts=0, fused_forward=None):
        """run-time schedule choices (``bpf_set_schedule``; 0 keeps one);
        ``fused_forward`` (None keeps it): ``set_fused_forward``"""
        self._check(self.L.bpf_set_schedule(self.h, int(newton_per_graph),
                                            int(tail_inverse), int(sweep_form),
                                            int(parts)),
                    "bpf_set_schedule")
        if fused_forward is not None:
            self.set_fused_forward(fused_forward)

    def set_fused_forward(self, max_active=32):
        """the Newton step's factorization forms the forward substitution
        (and the first solve skips its forward sweep) while at most
        ``max_active`` instances of a batch part are active; -1 never
        (``bpf_set_fused_forward``)"""
        self._check(self.L.bpf_set_fused_forward(self.h, int(max_active)),
                    "bpf_set_fused_forward")

    def set_factor_split(self, workgroups=0):
        """workgroups per instance for the factorization's subtrees while few
        instances are active (``bpf_set_factor_split``): 0 adaptive, 1 never,
        2 / 4 / 8 always"""
        self._check(self.L.bpf_set_factor_split(self.h, int(workgroups)),
                    "bpf_set_factor_split")

    def set_launch(self, lu=0, solve=0, mismatch=0):
        self._check(self.L.bpf_set_launch(self.h, lu, solve, mismatch),
                    "bpf_set_launch")
