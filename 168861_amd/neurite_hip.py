"""Python mirror of the Neurite solver interface over the MI355X C ABI.

Wraps ``lib/libneurite_hip.so`` (include/neurite_hip.h) with ctypes. The classes
mirror the reference's solver interface (inc/solver.h:12-24): ``Solver.update``,
``Solver.calculate``, ``Solver.snapshot``; ``Solver.run`` is the device-resident
time loop of neurite.cpp:224-255. There is no CPU fallback: every call goes to
the HIP library, and importing this module without the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libneurite_hip.so")

NRT_OK = 0
STATUS = {0: "ok", 1: "invalid", 2: "hip", 3: "nomem", 4: "topology", 5: "not_converged",
          6: "numeric", 7: "unsupported", 8: "admission", 9: "no_device"}
STATUS_CODES = {v: k for k, v in STATUS.items()}
EXPLICIT, IMPLICIT = 0, 1
FP64, FP32_REFINE = 0, 1
BICGSTAB, GMRES, DIRECT = 0, 1, 2
PRECOND_NONE, PRECOND_TREE = 0, 1
TOL_ABSOLUTE, TOL_RELATIVE = 0, 1
X0_PREVIOUS, X0_RHS, X0_ZERO = 0, 1, 2
TREE_BLOCK, TREE_GRID, TREE_LOCAL, TREE_LOCAL_LDS = 0, 1, 2, 3
# demangled name of the tree-solve kernel of each launch form (nrt_stats.tree_form)
TREE_KERNELS = {TREE_BLOCK: "k_implicit_tree<{T}, false>", TREE_GRID: "k_implicit_tree<{T}, true>",
                TREE_LOCAL: "k_implicit_tree_local<{T}, false>",
                TREE_LOCAL_LDS: "k_implicit_tree_local<{T}, true>"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint8)


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("layers_my", C.c_int32), ("rho_a", C.c_double),
                ("rho_m", C.c_double), ("per_m", C.c_double), ("thickness", C.c_double),
                ("rest_pot", C.c_double), ("thickness_my", C.c_double), ("per_my", C.c_double),
                ("ro_my", C.c_double), ("sigma_Na", C.c_double), ("sigma_K", C.c_double),
                ("sigma_L", C.c_double), ("dE_Na", C.c_double), ("dE_K", C.c_double),
                ("n_damage", C.c_double)]


class GeometryStruct(C.Structure):
    _fields_ = [("num_elem", C.c_int32), ("mother", _ip), ("dr", _ip), ("dl", _ip),
                ("branching", _up), ("diameter", _dp), ("dx0", _dp), ("epsilon", _dp),
                ("input", _dp), ("material", _ip), ("num_materials", C.c_int32),
                ("materials", C.POINTER(Material)), ("myelinate", C.c_int32),
                ("num_term", C.c_int32), ("terminals", _ip), ("E_MAX", C.c_double)]


class ProblemStruct(C.Structure):
    _fields_ = [("num_elem", C.c_int32), ("dT", C.c_double), ("potential0", _dp),
                ("potential1", _dp), ("dx", _dp), ("c_m", _dp), ("r_a", _dp), ("W", _dp),
                ("K", _dp), ("input", _dp), ("mother", _ip), ("dr", _ip), ("dl", _ip),
                ("branching", _up), ("num_term", C.c_int32), ("terminals", _ip),
                ("num_hh", C.c_int32), ("hh_elem", _ip), ("hh_m", _dp), ("hh_h", _dp),
                ("hh_n", _dp), ("E_Na", _dp), ("E_K", _dp), ("E_L", _dp), ("g_Na", _dp),
                ("g_K", _dp), ("G_L", _dp), ("ls_Na", _dp), ("ls_K", _dp), ("rest_pot", _dp)]


class SetupOutStruct(C.Structure):
    _fields_ = [("dx", _dp), ("c_m", _dp), ("r_a", _dp), ("r_m", _dp), ("W", _dp), ("K", _dp),
                ("crit_dT", _dp), ("rest", _dp), ("hh_elem", _ip), ("hh_m", _dp), ("hh_h", _dp),
                ("hh_n", _dp), ("E_Na", _dp), ("E_K", _dp), ("E_L", _dp), ("g_Na", _dp),
                ("g_K", _dp), ("G_L", _dp), ("ls_Na", _dp), ("ls_K", _dp), ("hh_rest", _dp),
                ("dT_crit", C.c_double), ("num_hh", C.c_int32)]


class ScenarioInfo(C.Structure):
    _fields_ = [("geometry", GeometryStruct), ("explicit_mode", C.c_int32),
                ("implicit_mode", C.c_int32), ("viewer", C.c_int32), ("mes_point", C.c_int32),
                ("tspan", C.c_double), ("xspan", C.c_double), ("dt_factor", C.c_double),
                ("nu", C.c_double), ("epsilon", C.c_double), ("E_MAX", C.c_double),
                ("init_time", C.c_double), ("end_input", C.c_double),
                ("input_current", C.c_double), ("dist_mes", C.c_double),
                ("epsilon_macro", C.c_double), ("out_path", C.c_char_p),
                ("log_path", C.c_char_p), ("builder_log", C.c_char_p),
                ("max_steps", C.c_int64), ("method", C.c_int32), ("precision", C.c_int32),
                ("restart", C.c_int32), ("tol", C.c_double), ("vout_path", C.c_char_p)]


class Options(C.Structure):
    _fields_ = [("kind", C.c_int32), ("precision", C.c_int32), ("krylov", C.c_int32),
                ("restart", C.c_int32), ("tol_mode", C.c_int32), ("x0_mode", C.c_int32),
                ("check_every", C.c_int32), ("max_iter", C.c_int32), ("device", C.c_int32),
                ("num_blocks", C.c_int32), ("tol", C.c_double), ("refine_tol", C.c_double),
                ("refine_max", C.c_int32), ("precond", C.c_int32), ("chunk", C.c_int32),
                ("chunk_hi", C.c_int32), ("batch_threads", C.c_int32),
                ("tree_groups", C.c_int32), ("reserved", C.c_int32 * 2)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("krylov_iters", C.c_int64), ("refine_sweeps", C.c_int64),
                ("last_iters", C.c_int32), ("last_status", C.c_int32),
                ("last_residual", C.c_double), ("last_rhs_norm", C.c_double),
                ("kernel_ms", C.c_double), ("kernel_launches", C.c_int64),
                ("matrix_nnz", C.c_int64), ("num_blocks", C.c_int32), ("restarts", C.c_int32),
                ("tree_form", C.c_int32), ("reserved", C.c_int32)]


class SlotInfo(C.Structure):
    _fields_ = [("N", C.c_int32), ("status", C.c_int32), ("steps_done", C.c_int64),
                ("num_steps", C.c_int64), ("inner_iters", C.c_int64),
                ("trace_samples", C.c_int64), ("nnz", C.c_int64)]


SIZE_CLASSES = ("N<4096", "4096<=N<16384", "N>=16384")  # batch slot size classes (rows)


class BatchTuneResult(C.Structure):
    _fields_ = [("threads", C.c_int32 * 3), ("chunk", C.c_int32 * 3),
                ("chunk_hi", C.c_int32 * 3), ("slots", C.c_int32 * 3), ("rows", C.c_int64 * 3),
                ("ns_per_row_step", C.c_double * 3), ("auto_ns_per_row_step", C.c_double * 3),
                ("trials", C.c_int32), ("reserved", C.c_int32), ("seconds", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("slot_steps", C.c_int64),
                ("compartment_steps", C.c_int64), ("kernel_ms", C.c_double),
                ("active_slots", C.c_int32), ("threads", C.c_int32),
                ("device_bytes", C.c_int64), ("class_threads", C.c_int32 * 3),
                ("reserved", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("steps_per_launch", C.c_int32), ("big_threshold", C.c_int32),
                ("max_big", C.c_int32), ("keep_results", C.c_int32), ("builders", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("N", C.c_int32),
                ("num_steps", C.c_int64), ("steps_done", C.c_int64), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("finish_s", C.c_double),
                ("trace_samples", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("running", C.c_int64), ("launches", C.c_int64),
                ("compartment_steps", C.c_int64), ("wall_s", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("batch_kernel_ms", C.c_double), ("launch_ms", C.c_double),
                ("admit_ms", C.c_double), ("wait_ms", C.c_double), ("retire_ms", C.c_double)]


class NrtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C 168861_amd)")
        L = C.CDLL(LIB_PATH)
        L.nrt_last_error.restype = C.c_char_p
        L.nrt_last_error.argtypes = [C.c_void_p]
        L.nrt_status_string.restype = C.c_char_p
        L.nrt_default_options.argtypes = [C.POINTER(Options)]
        L.nrt_setup.argtypes = [C.POINTER(GeometryStruct), C.c_int, C.POINTER(SetupOutStruct)]
        L.nrt_create.argtypes = [C.POINTER(ProblemStruct), C.POINTER(Options),
                                 C.POINTER(C.c_void_p)]
        L.nrt_update.argtypes = [C.c_void_p, C.c_int]
        L.nrt_calculate.argtypes = [C.c_void_p]
        L.nrt_snapshot.argtypes = [C.c_void_p, _dp]
        L.nrt_destroy.argtypes = [C.c_void_p]
        L.nrt_run.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int32,
                              C.c_int64, _dp, C.POINTER(C.c_int64)]
        L.nrt_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.nrt_sync.argtypes = [C.c_void_p]
        L.nrt_device_count.argtypes = [C.POINTER(C.c_int)]
        L.nrt_batch_create.argtypes = [C.POINTER(Options), C.POINTER(C.c_void_p)]
        L.nrt_batch_add.argtypes = [C.c_void_p, C.POINTER(ProblemStruct), C.c_int64, C.c_int64,
                                    C.c_int64, C.c_int32, C.c_int64, C.POINTER(C.c_int32)]
        L.nrt_batch_step.argtypes = [C.c_void_p, C.c_int64]
        L.nrt_batch_snapshot.argtypes = [C.c_void_p, C.c_int32, _dp]
        L.nrt_batch_trace.argtypes = [C.c_void_p, C.c_int32, _dp, C.c_int64,
                                      C.POINTER(C.c_int64)]
        L.nrt_batch_slot_info.argtypes = [C.c_void_p, C.c_int32, C.POINTER(SlotInfo)]
        L.nrt_batch_remove.argtypes = [C.c_void_p, C.c_int32]
        L.nrt_batch_get_stats.argtypes = [C.c_void_p, C.POINTER(BatchStats)]
        L.nrt_batch_set_shape.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32]
        L.nrt_batch_tune.argtypes = [C.c_void_p, C.POINTER(C.POINTER(ProblemStruct)), C.c_int32,
                                     C.c_int64, C.c_int64, C.POINTER(BatchTuneResult)]
        L.nrt_batch_last_error.argtypes = [C.c_void_p]
        L.nrt_batch_last_error.restype = C.c_char_p
        L.nrt_batch_destroy.argtypes = [C.c_void_p]
        L.nrt_default_server_config.argtypes = [C.POINTER(ServerConfig)]
        L.nrt_server_create.argtypes = [C.POINTER(Options), C.POINTER(ServerConfig),
                                        C.POINTER(C.c_void_p)]
        L.nrt_server_submit.argtypes = [C.c_void_p, C.POINTER(ProblemStruct), C.c_int64,
                                        C.c_int64, C.c_int64, C.c_int32, C.c_int64,
                                        C.POINTER(C.c_int64)]
        L.nrt_server_poll.argtypes = [C.c_void_p, C.POINTER(RequestResult), C.c_int32,
                                      C.POINTER(C.c_int32)]
        L.nrt_server_fetch.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, C.c_int64]
        L.nrt_server_wait_idle.argtypes = [C.c_void_p, C.c_double]
        L.nrt_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        L.nrt_server_last_error.argtypes = [C.c_void_p]
        L.nrt_server_last_error.restype = C.c_char_p
        L.nrt_server_destroy.argtypes = [C.c_void_p]
        L.nrt_scenario_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_char_p,
                                          C.POINTER(C.c_void_p)]
        L.nrt_scenario_get.argtypes = [C.c_void_p, C.POINTER(ScenarioInfo)]
        L.nrt_scenario_last_error.restype = C.c_char_p
        L.nrt_scenario_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(status, handle=None):
    if status != NRT_OK:
        msg = lib().nrt_last_error(handle).decode(errors="replace")
        raise NrtError(status, msg)


def device_count():
    n = C.c_int(0)
    _check(lib().nrt_device_count(C.byref(n)))
    return n.value


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


def default_options(**kw):
    o = Options()
    lib().nrt_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


# Per-element derived arrays (what the reference GPU constructors copy out of the
# discretization objects, solverEx-gpu.cpp:327-399).
PROBLEM_FIELDS_N = ["potential0", "potential1", "dx", "c_m", "r_a", "W", "K", "input"]
PROBLEM_FIELDS_H = ["hh_m", "hh_h", "hh_n", "E_Na", "E_K", "E_L", "g_Na", "g_K", "G_L",
                    "ls_Na", "ls_K", "rest_pot"]


class Problem:
    """Derived problem (numpy arrays), convertible to the C nrt_problem."""

    def __init__(self, dT, mother, dr, dl, branching, terminals, hh_elem, **arrays):
        self.dT = float(dT)
        self.mother = _i32(mother)
        self.dr = _i32(dr)
        self.dl = _i32(dl)
        self.branching = np.ascontiguousarray(branching, dtype=np.uint8)
        self.terminals = _i32(terminals)
        self.hh_elem = _i32(hh_elem)
        self.N = len(self.mother)
        for f in PROBLEM_FIELDS_N:
            v = arrays.get(f)
            if v is None:
                v = arrays["potential0"] if f == "potential1" else np.zeros(self.N)
            setattr(self, f, _f64(v))
        for f in PROBLEM_FIELDS_H:
            setattr(self, f, _f64(arrays[f]))
        self.extra = {k: v for k, v in arrays.items()
                      if k not in PROBLEM_FIELDS_N and k not in PROBLEM_FIELDS_H}

    @property
    def H(self):
        return len(self.hh_elem)

    def struct(self):
        s = ProblemStruct()
        s.num_elem = self.N
        s.dT = self.dT
        for f in PROBLEM_FIELDS_N:
            setattr(s, f, _ptr(getattr(self, f), _dp))
        s.mother, s.dr, s.dl = (_ptr(self.mother, _ip), _ptr(self.dr, _ip), _ptr(self.dl, _ip))
        s.branching = _ptr(self.branching, _up)
        s.num_term = len(self.terminals)
        s.terminals = _ptr(self.terminals, _ip)
        s.num_hh = self.H
        s.hh_elem = _ptr(self.hh_elem, _ip)
        for f in PROBLEM_FIELDS_H:
            setattr(s, f, _ptr(getattr(self, f), _dp))
        return s

    def with_dt(self, dT):
        p = Problem.__new__(Problem)
        p.__dict__.update(self.__dict__)
        p.dT = float(dT)
        return p


class Solver:
    """One simulation on the GPU: the reference's Solver (inc/solver.h:12-24)."""

    def __init__(self, problem, options=None, **kw):
        opts = options if options is not None else default_options(**kw)
        self._problem = problem
        self._s = problem.struct()
        h = C.c_void_p()
        _check(lib().nrt_create(C.byref(self._s), C.byref(opts), C.byref(h)))
        self.h = h
        self.N = problem.N

    def _c(self, st):
        _check(st, self.h)

    def update(self, input_active):
        self._c(lib().nrt_update(self.h, int(bool(input_active))))

    def calculate(self):
        self._c(lib().nrt_calculate(self.h))

    def snapshot(self):
        out = np.empty(self.N)
        self._c(lib().nrt_snapshot(self.h, out.ctypes.data_as(_dp)))
        return out

    def run(self, first_step, num_steps, active_first=-1, active_last=-2, trace_elem=0,
            interval=-1, check=True):
        """Device-resident time loop; returns the trace samples (numpy)."""
        cap = num_steps // (interval + 1) if interval >= 0 else 0
        trace = np.empty(max(cap, 1))
        ns = C.c_int64(0)
        st = lib().nrt_run(self.h, first_step, num_steps, active_first, active_last, trace_elem,
                           interval, trace.ctypes.data_as(_dp) if cap else None, C.byref(ns))
        if check:
            self._c(st)
        self.last_status = st
        return trace[:ns.value]

    def stats(self):
        s = Stats()
        self._c(lib().nrt_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def sync(self):
        self._c(lib().nrt_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().nrt_destroy(self.h)
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


class Batch:
    """Many independent neurites advanced together (nrt_batch_*): each slot is the
    problem a Solver would run alone; slots join and leave between steps."""

    def __init__(self, options=None, **kw):
        opts = options if options is not None else default_options(**kw)
        b = C.c_void_p()
        st = lib().nrt_batch_create(C.byref(opts), C.byref(b))
        if st != 0:
            raise NrtError(st, lib().nrt_last_error(None).decode() or "batch creation failed")
        self.b = b
        self._keep = {}

    def _c(self, st):
        if st != 0:
            raise NrtError(st, lib().nrt_batch_last_error(self.b).decode())

    def add(self, problem, num_steps, a0=1, a1=0, trace_elem=0, trace_every=-1):
        s = problem.struct()
        slot = C.c_int32(-1)
        self._c(lib().nrt_batch_add(self.b, C.byref(s), num_steps, a0, a1, trace_elem,
                                    trace_every, C.byref(slot)))
        self._keep[slot.value] = problem.N
        return slot.value

    def step(self, k):
        self._c(lib().nrt_batch_step(self.b, k))

    def set_shapes(self, classes):
        """Launch shapes per size class, as tune() returns them ({class name:
        {"threads", "chunk", "chunk_hi"}}; "auto" chunks are the auto rule)."""
        for c, name in enumerate(SIZE_CLASSES):
            if name in classes:
                sh = classes[name]
                ch = 0 if sh["chunk"] == "auto" else int(sh["chunk"])
                chh = 0 if sh["chunk_hi"] == "auto" else int(sh["chunk_hi"])
                self._c(lib().nrt_batch_set_shape(self.b, c, int(sh["threads"]), ch, chh))

    def tune(self, problems, steps=10, rows_per_class=4_000_000):
        """Launch shapes per size class (nrt_batch_tune) from a sample of problems:
        returns, per class (rows < 4096, < 16384, the rest), the chosen workgroup
        size and chunk lengths with their kernel time per compartment-step."""
        structs = [p.struct() for p in problems]
        arr = (C.POINTER(ProblemStruct) * max(1, len(structs)))(
            *[C.pointer(x) for x in structs])
        r = BatchTuneResult()
        self._c(lib().nrt_batch_tune(self.b, arr, len(structs), steps, rows_per_class,
                                     C.byref(r)))
        names = SIZE_CLASSES
        out = {"trials": r.trials, "seconds": r.seconds, "classes": {}}
        for c, name in enumerate(names):
            if r.slots[c] == 0:
                continue
            out["classes"][name] = {
                "threads": r.threads[c], "chunk": r.chunk[c] or "auto",
                "chunk_hi": r.chunk_hi[c] or "auto", "sample_trees": r.slots[c],
                "sample_rows": r.rows[c], "ns_per_compartment_step": r.ns_per_row_step[c],
                "auto_256_ns_per_compartment_step": r.auto_ns_per_row_step[c]}
        return out

    def snapshot(self, slot):
        out = np.empty(self._keep[slot])
        self._c(lib().nrt_batch_snapshot(self.b, slot, out.ctypes.data_as(_dp)))
        return out

    def trace(self, slot):
        n = C.c_int64(0)
        self._c(lib().nrt_batch_trace(self.b, slot, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1))
        self._c(lib().nrt_batch_trace(self.b, slot, out.ctypes.data_as(_dp), n.value,
                                      C.byref(n)))
        return out[:n.value]

    def info(self, slot):
        i = SlotInfo()
        self._c(lib().nrt_batch_slot_info(self.b, slot, C.byref(i)))
        return {f: getattr(i, f) for f, _ in SlotInfo._fields_}

    def remove(self, slot):
        self._c(lib().nrt_batch_remove(self.b, slot))
        self._keep.pop(slot, None)

    def stats(self):
        s = BatchStats()
        self._c(lib().nrt_batch_get_stats(self.b, C.byref(s)))
        out = {f: getattr(s, f) for f, _ in BatchStats._fields_ if f != "reserved"}
        out["class_threads"] = list(s.class_threads)
        return out

    def close(self):
        if getattr(self, "b", None):
            lib().nrt_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Request server (nrt_server_*): submitted simulations are admitted by free
    device memory and advanced together (continuous batching) by a native loop."""

    def __init__(self, options=None, steps_per_launch=16, big_threshold=1 << 30, max_big=4,
                 keep_results=True, builders=4, **kw):
        opts = options if options is not None else default_options(**kw)
        cfg = ServerConfig()
        lib().nrt_default_server_config(C.byref(cfg))
        cfg.steps_per_launch = steps_per_launch
        cfg.big_threshold = big_threshold
        cfg.max_big = max_big
        cfg.keep_results = int(bool(keep_results))
        cfg.builders = builders
        s = C.c_void_p()
        st = lib().nrt_server_create(C.byref(opts), C.byref(cfg), C.byref(s))
        if st != 0:
            raise NrtError(st, lib().nrt_last_error(None).decode() or "server creation failed")
        self.s = s
        self._req = {}

    def _c(self, st):
        if st != 0:
            raise NrtError(st, lib().nrt_server_last_error(self.s).decode())

    def submit(self, problem, num_steps, a0=1, a1=0, trace_elem=0, trace_every=-1):
        ps = problem.struct()
        rid = C.c_int64(0)
        self._c(lib().nrt_server_submit(self.s, C.byref(ps), num_steps, a0, a1, trace_elem,
                                        trace_every, C.byref(rid)))
        self._req[rid.value] = (problem.N, num_steps // (trace_every + 1)
                                if trace_every >= 0 else 0)
        return rid.value

    def poll(self, capacity=4096):
        buf = (RequestResult * capacity)()
        n = C.c_int32(0)
        self._c(lib().nrt_server_poll(self.s, buf, capacity, C.byref(n)))
        return [{f: getattr(buf[i], f) for f, _ in RequestResult._fields_}
                for i in range(n.value)]

    def fetch(self, rid):
        """(status, final V, trace) of a completed request; releases them."""
        N, ns = self._req.pop(rid)
        v = np.empty(N)
        tr = np.empty(max(ns, 1))
        st = lib().nrt_server_fetch(self.s, rid, v.ctypes.data_as(_dp),
                                    tr.ctypes.data_as(_dp), ns)
        if st not in STATUS:
            self._c(st)
        return st, v, tr[:ns]

    def wait_idle(self, timeout=-1.0):
        self._c(lib().nrt_server_wait_idle(self.s, timeout))

    def stats(self):
        s = ServerStats()
        self._c(lib().nrt_server_get_stats(self.s, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "s", None):
            lib().nrt_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Geometry:
    """Raw geometry (what a space.cpp builder produces), numpy arrays."""

    def __init__(self, mother, dr, dl, branching, diameter, dx0, epsilon, input, material,
                 materials, terminals, myelinate, E_MAX=1.0):
        self.mother, self.dr, self.dl = _i32(mother), _i32(dr), _i32(dl)
        self.branching = np.ascontiguousarray(branching, dtype=np.uint8)
        self.diameter, self.dx0 = _f64(diameter), _f64(dx0)
        self.epsilon, self.input = _f64(epsilon), _f64(input)
        self.material = _i32(material)
        self.materials = list(materials)
        self.terminals = _i32(terminals)
        self.myelinate = int(myelinate)
        self.E_MAX = float(E_MAX)
        self.N = len(self.mother)

    def struct(self):
        g = GeometryStruct()
        g.num_elem = self.N
        g.mother, g.dr, g.dl = _ptr(self.mother, _ip), _ptr(self.dr, _ip), _ptr(self.dl, _ip)
        g.branching = _ptr(self.branching, _up)
        g.diameter, g.dx0 = _ptr(self.diameter, _dp), _ptr(self.dx0, _dp)
        g.epsilon, g.input = _ptr(self.epsilon, _dp), _ptr(self.input, _dp)
        g.material = _ptr(self.material, _ip)
        self._mats = (Material * len(self.materials))(*self.materials)
        g.num_materials = len(self.materials)
        g.materials = C.cast(self._mats, C.POINTER(Material))
        g.myelinate = self.myelinate
        g.num_term = len(self.terminals)
        g.terminals = _ptr(self.terminals, _ip)
        g.E_MAX = self.E_MAX
        return g


def setup_out_buffers(N, H):
    bufs = {k: np.zeros(N) for k in ["dx", "c_m", "r_a", "r_m", "W", "K", "crit_dT", "rest"]}
    bufs.update({k: np.zeros(max(H, 1)) for k in ["hh_m", "hh_h", "hh_n", "E_Na", "E_K", "E_L",
                                                  "g_Na", "g_K", "G_L", "ls_Na", "ls_K",
                                                  "hh_rest"]})
    bufs["hh_elem"] = np.zeros(max(H, 1), dtype=np.int32)
    s = SetupOutStruct()
    for k, v in bufs.items():
        setattr(s, k, _ptr(v, _ip if v.dtype == np.int32 else _dp))
    return s, bufs


def count_hh(geom):
    kinds = np.array([m.kind for m in geom.materials])
    return int((kinds[geom.material] == 1).sum())


def setup(geom, dT=None, dt_factor=1.0, device=-1):
    """Device setup chain (init.cpp:138-298): geometry -> Problem.

    dT defaults to dt_factor * (critical time step), as neurite.cpp:171-173.
    Returns (problem, dT_crit).
    """
    H = count_hh(geom)
    s, bufs = setup_out_buffers(geom.N, H)
    g = geom.struct()
    _check(lib().nrt_setup(C.byref(g), device, C.byref(s)))
    dT_crit = s.dT_crit
    if dT is None:
        dT = dT_crit * dt_factor
    rest = bufs["rest"]
    prob = Problem(dT, geom.mother, geom.dr, geom.dl, geom.branching, geom.terminals,
                   bufs["hh_elem"][:H], potential0=rest, potential1=rest, dx=bufs["dx"],
                   c_m=bufs["c_m"], r_a=bufs["r_a"], W=bufs["W"], K=bufs["K"], input=geom.input,
                   hh_m=bufs["hh_m"][:H], hh_h=bufs["hh_h"][:H], hh_n=bufs["hh_n"][:H],
                   E_Na=bufs["E_Na"][:H], E_K=bufs["E_K"][:H], E_L=bufs["E_L"][:H],
                   g_Na=bufs["g_Na"][:H], g_K=bufs["g_K"][:H], G_L=bufs["G_L"][:H],
                   ls_Na=bufs["ls_Na"][:H], ls_K=bufs["ls_K"][:H], rest_pot=bufs["hh_rest"][:H],
                   r_m=bufs["r_m"], crit_dT=bufs["crit_dT"])
    return prob, dT_crit


def active_range(dT, num_steps_T, init_time, end_input):
    """Stimulus window of neurite.cpp:227-230 as an inclusive step range
    (the same floating-point comparisons, evaluated once on the host)."""
    i = np.arange(1, max(num_steps_T - 1, 1), dtype=np.int64)
    t = i * float(dT)
    on = (t >= init_time) & (t <= end_input)
    if not on.any():
        return -1, -2
    idx = i[on]
    return int(idx[0]), int(idx[-1])


def scenario(args, simulation=None, working_directory=None):
    """The reference's pre-solver chain on a command line (nrt_scenario_*):
    tokens, the builder named by `simulation`, the mechanical model, the
    measurement point. Returns a dict: `geometry` (Geometry) and the run values."""
    argv = [b"neurite_hip"] + [a.encode() for a in args]
    arr = (C.c_char_p * len(argv))(*argv)
    h = C.c_void_p()
    L = lib()
    st = L.nrt_scenario_create(len(argv), arr,
                               simulation.encode() if simulation else None,
                               working_directory.encode() if working_directory else None,
                               C.byref(h))
    if st != 0:
        raise NrtError(st, L.nrt_scenario_last_error().decode())
    try:
        info = ScenarioInfo()
        _check(L.nrt_scenario_get(h, C.byref(info)))
        g = info.geometry
        N = g.num_elem

        def arr_(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        mats = [Material.from_buffer_copy(g.materials[k]) for k in range(g.num_materials)]
        geom = Geometry(arr_(g.mother, N, np.int32), arr_(g.dr, N, np.int32),
                        arr_(g.dl, N, np.int32), arr_(g.branching, N, np.uint8),
                        arr_(g.diameter, N, np.float64), arr_(g.dx0, N, np.float64),
                        arr_(g.epsilon, N, np.float64), arr_(g.input, N, np.float64),
                        arr_(g.material, N, np.int32), mats,
                        arr_(g.terminals, g.num_term, np.int32), g.myelinate, g.E_MAX)
        out = {f: getattr(info, f) for f, _ in ScenarioInfo._fields_ if f != "geometry"}
        for k in ("out_path", "log_path", "builder_log", "vout_path"):
            out[k] = out[k].decode() if out[k] is not None else None
        out["geometry"] = geom
        return out
    finally:
        L.nrt_scenario_destroy(h)
