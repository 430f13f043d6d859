"""feyntrop_amd -- MI355X tropical Monte Carlo sampler for Feynman integrals.

Python binding of the C ABI in include/feyntrop_amd.h (libfeyntrop_amd.so,
built in-tree by feyntrop_amd/csrc/Makefile).  The library is the only
compute path: if it is missing this module raises on import of the engine.

Reference operator interface mirrored (michibo/feyntrop):
  Engine.run_json(text)      -- the ./feyntrop stdin/stdout JSON contract
                                 (srcs/feyntrop.cpp:20-281)
  Engine.prepare(problem)    -- graph + checks + regime + J-table
                                 (feyntrop.cpp:45-239)
  Engine.integrate(p, N, s)  -- feynman_integral_estimate
                                 (feynman_integral.hpp:30-373) -> raw moments
  Problem.finalize(moments)  -- stats avg/acc (stats.hpp:86-104)
"""

import ctypes as C
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfeyntrop_amd.so")
BIN_PATH = os.path.join(_HERE, "bin", "feyntrop")
KCACHE = os.path.join(_HERE, "kcache")

FTR_MAX_K = 16


class ftr_input(C.Structure):
    _fields_ = [("E", C.c_int), ("V", C.c_int), ("edges", C.POINTER(C.c_int32)),
                ("nu", C.POINTER(C.c_double)), ("P", C.POINTER(C.c_double)),
                ("m2", C.POINTER(C.c_double)), ("D", C.c_double), ("K", C.c_int),
                ("lambda_", C.c_double)]


class ftr_info(C.Structure):
    _fields_ = [("E", C.c_int), ("V", C.c_int), ("L", C.c_int), ("K", C.c_int),
                ("W", C.c_double), ("IGtr", C.c_double),
                ("euclidean", C.c_int), ("pseudo_euclidean", C.c_int), ("generic", C.c_int),
                ("gp_property", C.c_int), ("lambda_eff", C.c_double), ("path", C.c_int),
                ("seconds_preprocessing", C.c_double)]


class ftr_moments(C.Structure):
    _fields_ = [("K", C.c_int),
                ("n", (C.c_double * 2) * FTR_MAX_K),
                ("s1", (C.c_double * 2) * FTR_MAX_K),
                ("s2", (C.c_double * 2) * FTR_MAX_K),
                ("dropped", C.c_uint64), ("kernel_ms", C.c_double), ("seconds", C.c_double),
                ("lanes", C.c_uint64), ("samples_per_lane", C.c_uint64)]

    def arrays(self):
        K = self.K
        n = np.array([[self.n[k][q] for q in range(2)] for k in range(K)])
        s1 = np.array([[self.s1[k][q] for q in range(2)] for k in range(K)])
        s2 = np.array([[self.s2[k][q] for q in range(2)] for k in range(K)])
        return n, s1, s2


class ftr_opts(C.Structure):
    _fields_ = [("block", C.c_int), ("table_mode", C.c_int), ("lanes", C.c_uint64)]


_lib = None


def lib():
    """Load libfeyntrop_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"feyntrop_amd: {LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.ftr_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ftr_destroy.argtypes = [vp]
    L.ftr_last_error.restype = C.c_char_p
    L.ftr_last_error.argtypes = [vp]
    L.ftr_prepare.argtypes = [vp, C.POINTER(ftr_input), C.POINTER(vp)]
    L.ftr_problem_free.argtypes = [vp]
    L.ftr_problem_info.argtypes = [vp, C.POINTER(ftr_info)]
    L.ftr_jtable.argtypes = [vp, vp, vp, vp, vp]
    L.ftr_integrate.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(ftr_opts), C.POINTER(ftr_moments)]
    L.ftr_precompile_batch.argtypes = [C.c_int, C.POINTER(ftr_input), C.c_char_p, C.c_int]
    L.ftr_integrate_noalloc.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(ftr_opts),
                                        C.POINTER(ftr_moments)]
    L.ftr_finalize.argtypes = [vp, C.POINTER(ftr_moments), vp]
    L.ftr_sample.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp]
    L.ftr_replay.argtypes = [vp, vp, C.c_uint64, vp, vp, vp, vp]
    L.ftr_run_json.argtypes = [vp, C.c_char_p, C.POINTER(vp), C.POINTER(vp)]
    L.ftr_free_string.argtypes = [vp]
    L.ftr_integrate_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64), C.POINTER(ftr_moments)]
    L.ftr_set_cache_dir.argtypes = [C.c_char_p]
    L.ftr_autotune.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.ftr_set_variant.argtypes = [vp, vp, C.c_int, C.c_int]
    L.ftr_get_variant.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ftr_set_lanes_per_sample.argtypes = [vp, vp, C.c_int]
    L.ftr_get_lanes_per_sample.argtypes = [vp, C.POINTER(C.c_int)]
    L.ftr_precompile_variant.argtypes = [C.POINTER(ftr_input), C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                         C.c_int]
    L.ftr_precompile.argtypes = [C.POINTER(ftr_input), C.c_char_p, C.c_int]
    L.ftr_set_jtable_mode.argtypes = [vp, C.c_int]
    L.ftr_batch_graph_replays.argtypes = [vp]
    L.ftr_prepare_batch.argtypes = [vp, C.c_int, C.POINTER(ftr_input), C.POINTER(vp), C.POINTER(C.c_int)]
    L.ftr_batch_error.restype = C.c_char_p
    L.ftr_batch_error.argtypes = [vp, C.c_int]
    L.ftr_set_specialized.argtypes = [vp, vp, C.c_int]
    L.ftr_use_cached_specialization.argtypes = [vp, vp]
    L.ftr_get_specialized.argtypes = [vp, C.POINTER(C.c_int)]
    L.ftr_precompile_specialized.argtypes = [C.POINTER(ftr_input), C.c_char_p, C.c_int]
    L.ftr_precompile_learned.argtypes = [C.POINTER(ftr_input), C.c_char_p, C.c_int]
    L.ftr_learner_create.argtypes = [vp, vp, C.c_uint64, C.POINTER(vp)]
    L.ftr_learner_free.argtypes = [vp]
    L.ftr_learner_params.argtypes = [vp, vp, vp]
    L.ftr_learner_forward.argtypes = [vp, vp, vp]
    L.ftr_learner_backward.argtypes = [vp, vp, vp, vp]
    L.ftr_learner_train.argtypes = [vp, vp, C.c_int, C.c_uint64, C.c_uint64, C.c_double, vp]
    L.ftr_learner_tables.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.ftr_learner_integrate.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(ftr_opts),
                                        C.POINTER(ftr_moments)]
    L.ftr_learner_sample.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp, vp]
    L.ftr_learner_set_objective.argtypes = [vp, C.c_int]
    L.ftr_learner_statistics.argtypes = [vp, vp, C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp, vp, vp]
    L.ftr_learner_select.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


class _Input:
    """Keeps the numpy buffers of an ftr_input alive."""

    def __init__(self, prob):
        g = prob["graph"]
        self.edges = np.array([[int(e[0][0]), int(e[0][1])] for e in g], dtype=np.int32).reshape(-1)
        self.nu = np.array([float(e[1]) for e in g], dtype=np.float64)
        self.P = np.ascontiguousarray(np.array(prob["scalarproducts"], dtype=np.float64).reshape(-1))
        self.m2 = np.array([float(m) for m in prob["masses_sqr"]], dtype=np.float64)
        V = int(round(len(self.P) ** 0.5))
        self.s = ftr_input(len(self.nu), V,
                           self.edges.ctypes.data_as(C.POINTER(C.c_int32)),
                           self.nu.ctypes.data_as(C.POINTER(C.c_double)),
                           self.P.ctypes.data_as(C.POINTER(C.c_double)),
                           self.m2.ctypes.data_as(C.POINTER(C.c_double)),
                           float(prob["dimension"]), int(prob["num_eps_terms"]), float(prob["lambda"]))


def _field_dtype(ctype):
    if ctype in (C.c_int, C.c_int32):
        return np.int32
    if ctype is C.c_double:
        return np.float64
    return np.uint64   # pointers


# ftr_input as a numpy record, field for field (offsets and size from ctypes)
_INPUT_DTYPE = np.dtype({"names": [f for f, _ in ftr_input._fields_],
                         "formats": [_field_dtype(t) for _, t in ftr_input._fields_],
                         "offsets": [getattr(ftr_input, f).offset for f, _ in ftr_input._fields_],
                         "itemsize": C.sizeof(ftr_input)})


class _BatchInputs:
    """The ftr_input records of many problems at once: every field of every
    problem in one flat array per field, the records filled column-wise (a
    4096-point scan builds in milliseconds instead of a Python object per
    problem)."""

    def __init__(self, probs):
        from itertools import chain
        n = len(probs)
        graphs = [pr["graph"] for pr in probs]
        E = np.fromiter((len(g) for g in graphs), np.int64, n)
        V = np.fromiter((len(pr["scalarproducts"]) for pr in probs), np.int64, n)
        self.edges = np.fromiter(chain.from_iterable((e[0][0], e[0][1]) for g in graphs for e in g), np.int32,
                                 2 * int(E.sum()))
        self.nu = np.fromiter((e[1] for g in graphs for e in g), np.float64, int(E.sum()))
        self.P = np.fromiter(chain.from_iterable(chain.from_iterable(pr["scalarproducts"] for pr in probs)),
                             np.float64, int((V * V).sum()))
        self.m2 = np.fromiter(chain.from_iterable(pr["masses_sqr"] for pr in probs), np.float64, int(E.sum()))
        oe = np.concatenate([[0], np.cumsum(E)[:-1]])
        op = np.concatenate([[0], np.cumsum(V * V)[:-1]])
        self.rec = np.zeros(n, _INPUT_DTYPE)
        self.rec["E"], self.rec["V"] = E, V
        self.rec["edges"] = self.edges.ctypes.data + 8 * oe      # 2 int32 per edge
        self.rec["nu"] = self.nu.ctypes.data + 8 * oe
        self.rec["P"] = self.P.ctypes.data + 8 * op
        self.rec["m2"] = self.m2.ctypes.data + 8 * oe
        self.rec["D"] = np.fromiter((pr["dimension"] for pr in probs), np.float64, n)
        self.rec["K"] = np.fromiter((pr["num_eps_terms"] for pr in probs), np.int32, n)
        self.rec["lambda_"] = np.fromiter((pr["lambda"] for pr in probs), np.float64, n)

    def pointer(self):
        return self.rec.ctypes.data_as(C.POINTER(ftr_input))


class FeyntropError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def precompile(prob):
    """Specialise and cache the kernels of a problem (no GPU needed)."""
    inp = _Input(prob)
    err = C.create_string_buffer(8192)
    rc = lib().ftr_precompile(C.byref(inp.s), err, 8192)
    if rc:
        raise FeyntropError(rc, err.value.decode(errors="replace"))


def precompile_batch(probs):
    """Compile the bucket kernels of a batched integration of these problems
    (those the preprocessing accepts) into the kernel cache (no GPU needed)."""
    inps = _BatchInputs(probs)
    err = C.create_string_buffer(8192)
    rc = lib().ftr_precompile_batch(len(probs), inps.pointer(), err, 8192)
    if rc:
        raise FeyntropError(rc, err.value.decode(errors="replace"))


def precompile_learned(prob):
    """Specialise and cache the learned-proposal kernels of a problem."""
    inp = _Input(prob)
    err = C.create_string_buffer(8192)
    rc = lib().ftr_precompile_learned(C.byref(inp.s), err, 8192)
    if rc:
        raise FeyntropError(rc, err.value.decode(errors="replace"))


def precompile_specialized(prob):
    """Compile the kernel specialised to the problem's kinematic point."""
    inp = _Input(prob)
    err = C.create_string_buffer(8192)
    rc = lib().ftr_precompile_specialized(C.byref(inp.s), err, 8192)
    if rc:
        raise FeyntropError(rc, err.value.decode(errors="replace"))


def precompile_variant(prob, table_mode, block, lps=1, specialized=False):
    """Compile one kernel variant (table placement, block size, lanes per
    sample; optionally specialised to the kinematic point) into the cache."""
    inp = _Input(prob)
    err = C.create_string_buffer(8192)
    rc = lib().ftr_precompile_variant(C.byref(inp.s), int(table_mode), int(block), int(lps),
                                      1 if specialized else 0, err, 8192)
    if rc:
        raise FeyntropError(rc, err.value.decode(errors="replace"))


def host_prepare(prob):
    """Host-only preparation (no GPU): validation, regime flags, J-table."""
    inp = _Input(prob)
    out = C.c_void_p()
    rc = lib().ftr_prepare(None, C.byref(inp.s), C.byref(out))
    if rc:
        raise FeyntropError(rc, lib().ftr_last_error(None).decode(errors="replace"))
    return Problem(None, out, prob)


def host_path(prob):
    """Kernel path of a problem (0 real, 1 real psi-only, 2 deformed) or None
    if the preprocessing rejects it."""
    try:
        return host_prepare(prob).info.path
    except FeyntropError:
        return None


def host_accepts(prob):
    return host_path(prob) is not None


class Problem:
    def __init__(self, engine, handle, prob):
        self.engine = engine
        self.h = handle
        self.prob = prob
        info = ftr_info()
        lib().ftr_problem_info(self.h, C.byref(info))
        self.info = info
        self.E, self.V, self.K = info.E, info.V, info.K

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ftr_problem_free(self.h)
            self.h = None

    def jtable(self):
        n = 1 << self.E
        Jr, r = np.zeros(n), np.zeros(n)
        mm, L = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        lib().ftr_jtable(self.h, Jr.ctypes.data, r.ctypes.data, mm.ctypes.data, L.ctypes.data)
        return Jr, r, mm, L

    def finalize(self, m):
        res = np.zeros(self.K * 4)
        lib().ftr_finalize(self.h, C.byref(m), res.ctypes.data)
        return res.reshape(self.K, 2, 2)


class Engine:
    """One HIP device context (stream, kernel modules, scratch)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().ftr_create(device, C.byref(h))
        if rc:
            raise FeyntropError(rc, "ftr_create failed (no usable HIP device?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.ftr_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def __del__(self):
        self.close()

    def _err(self, rc):
        return FeyntropError(rc, lib().ftr_last_error(self.h).decode(errors="replace"))

    def prepare(self, prob):
        inp = _Input(prob)
        out = C.c_void_p()
        rc = lib().ftr_prepare(self.h, C.byref(inp.s), C.byref(out))
        if rc:
            raise self._err(rc)
        return Problem(self, out, prob)

    def batch_graph_replays(self):
        return lib().ftr_batch_graph_replays(self.h)

    def prepare_batch(self, probs):
        """Prepare many problems (host phases in parallel).  Returns a list with
        a Problem, or a FeyntropError for a rejected problem, per input."""
        n = len(probs)
        inps = _BatchInputs(probs)
        outs = (C.c_void_p * n)()
        st = (C.c_int * n)()
        rc = lib().ftr_prepare_batch(self.h, n, inps.pointer(), outs, st)
        if rc:
            raise self._err(rc)
        res = []
        for i in range(n):
            if st[i]:
                res.append(FeyntropError(st[i], lib().ftr_batch_error(self.h, i).decode(errors="replace")))
            else:
                res.append(Problem(self, C.c_void_p(outs[i]), probs[i]))
        return res

    def integrate(self, p, N, seed=0, lanes=0):
        m = ftr_moments()
        opts = ftr_opts(0, -1, int(lanes))
        rc = lib().ftr_integrate(self.h, p.h, int(N), int(seed), C.byref(opts), C.byref(m))
        if rc:
            raise self._err(rc)
        return m

    def integrate_noalloc(self, p, N, seed=0, lanes=0):
        """ftr_integrate_noalloc: the same integration with no device
        allocation (scratch in a fixed per-GPU device array)."""
        m = ftr_moments()
        opts = ftr_opts(0, -1, int(lanes))
        rc = lib().ftr_integrate_noalloc(self.h, p.h, int(N), int(seed), C.byref(opts), C.byref(m))
        if rc:
            raise self._err(rc)
        return m

    def integrate_batch(self, probs, Ns, seeds):
        n = len(probs)
        hs = (C.c_void_p * n)(*[p.h for p in probs])
        Na = (C.c_uint64 * n)(*[int(x) for x in Ns])
        Sa = (C.c_uint64 * n)(*[int(x) for x in seeds])
        out = (ftr_moments * n)()
        rc = lib().ftr_integrate_batch(self.h, n, hs, Na, Sa, out)
        if rc:
            raise self._err(rc)
        return list(out)

    def sample(self, p, first, n, seed, X_ptr, px_ptr):
        """Device pointers (e.g. torch tensor .data_ptr()) of n*E and n*2 doubles."""
        rc = lib().ftr_sample(self.h, p.h, int(first), int(n), int(seed), C.c_void_p(X_ptr), C.c_void_p(px_ptr))
        if rc:
            raise self._err(rc)

    def replay(self, p, n, X_ptr, px_ptr, out_ptr, drop_ptr):
        rc = lib().ftr_replay(self.h, p.h, int(n), C.c_void_p(X_ptr), C.c_void_p(px_ptr),
                              C.c_void_p(out_ptr), C.c_void_p(drop_ptr))
        if rc:
            raise self._err(rc)

    def autotune(self, p, n_probe=2_000_000):
        tm, blk, ms = C.c_int(), C.c_int(), C.c_double()
        rc = lib().ftr_autotune(self.h, p.h, int(n_probe), C.byref(tm), C.byref(blk), C.byref(ms))
        if rc:
            raise self._err(rc)
        return tm.value, blk.value, ms.value

    def set_variant(self, p, table_mode, block):
        rc = lib().ftr_set_variant(self.h, p.h, int(table_mode), int(block))
        if rc:
            raise self._err(rc)

    def set_lanes_per_sample(self, p, lps):
        """1: one sample per lane; 2 (deformed path): a lane pair per sample."""
        rc = lib().ftr_set_lanes_per_sample(self.h, p.h, int(lps))
        if rc:
            raise self._err(rc)

    def lanes_per_sample(self, p):
        v = C.c_int()
        lib().ftr_get_lanes_per_sample(p.h, C.byref(v))
        return v.value

    def set_specialized(self, p, on=True):
        """Specialise p's integration kernel to its kinematic point."""
        rc = lib().ftr_set_specialized(self.h, p.h, 1 if on else 0)
        if rc:
            raise self._err(rc)

    def use_cached_specialization(self, p):
        """The feyntrop binary's kernel choice: p's specialised kernel when it
        is already cached (FEYNTROP_AMD_SPECIALIZE=1 always, =0 never);
        returns whether p now runs it."""
        rc = lib().ftr_use_cached_specialization(self.h, p.h)
        if rc:
            raise self._err(rc)
        return self.specialized(p)

    def specialized(self, p):
        on = C.c_int()
        lib().ftr_get_specialized(p.h, C.byref(on))
        return bool(on.value)

    def set_jtable_mode(self, mode):
        """J-table construction: 0 auto, 1 host, 2 device."""
        rc = lib().ftr_set_jtable_mode(self.h, int(mode))
        if rc:
            raise self._err(rc)

    def variant(self, p):
        tm, blk = C.c_int(), C.c_int()
        lib().ftr_get_variant(p.h, C.byref(tm), C.byref(blk))
        return tm.value, blk.value

    def run_json(self, text):
        out, log = C.c_void_p(), C.c_void_p()
        rc = lib().ftr_run_json(self.h, text.encode(), C.byref(out), C.byref(log))
        o = C.string_at(out.value).decode() if out.value else None
        lg = C.string_at(log.value).decode() if log.value else ""
        lib().ftr_free_string(out)
        lib().ftr_free_string(log)
        return rc, o, lg


LEARN_HIDDEN = 128
LEARN_NPARAMS = 128 * 32 + 128 + 2 * (128 * 128 + 128) + 32 * 128 + 32


class Learner:
    """Learned proposal of one prepared problem (include/feyntrop_amd.h,
    ftr_learner_*).  Parameters: FP32 vector of LEARN_NPARAMS values
    (W1[128x32] b1 W2[128x128] b2 W3[128x128] b3 W4[32x128] b4)."""

    def __init__(self, engine, p, init_seed=0):
        self.engine, self.p = engine, p
        h = C.c_void_p()
        rc = lib().ftr_learner_create(engine.h, p.h, int(init_seed), C.byref(h))
        if rc:
            raise engine._err(rc)
        self.h = h
        self.E = p.E
        self.rows = max(32, 1 << p.E)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ftr_learner_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise self.engine._err(rc)

    def params(self):
        out = np.zeros(LEARN_NPARAMS, np.float32)
        self._chk(lib().ftr_learner_params(self.h, out.ctypes.data, None))
        return out

    def set_params(self, par):
        par = np.ascontiguousarray(par, dtype=np.float32)
        assert par.size == LEARN_NPARAMS
        self._chk(lib().ftr_learner_params(self.h, None, par.ctypes.data))

    def forward(self):
        Y = np.zeros((self.rows, 32), np.float32)
        self._chk(lib().ftr_learner_forward(self.engine.h, self.h, Y.ctypes.data))
        return Y

    def backward(self, dY):
        dY = np.ascontiguousarray(dY, dtype=np.float32)
        assert dY.shape == (self.rows, 32)
        g = np.zeros(LEARN_NPARAMS, np.float32)
        self._chk(lib().ftr_learner_backward(self.engine.h, self.h, dY.ctypes.data, g.ctypes.data))
        return g

    def train(self, steps, samples, seed=1, lr=3e-3):
        """Returns per step (relative second moment of the eps^0 real part, kernel ms)."""
        st = np.zeros((max(steps, 1), 2))
        self._chk(lib().ftr_learner_train(self.engine.h, self.h, int(steps), int(samples), int(seed),
                                          float(lr), st.ctypes.data))
        return st[:steps]

    def statistics(self, samples, seed):
        """One training step's record and forward-KL statistics, without the
        update: dict of c (N), ly ((E-1) x N), path (N), gC (2^E x E), gS
        (2^E x E), gsum (2) and dY (rows x 32)."""
        n, E, N = 1 << self.E, self.E, int(samples)
        out = dict(c=np.zeros(N), ly=np.zeros((max(E - 1, 1), N)), path=np.zeros(N, np.uint64),
                   gC=np.zeros((n, E)), gS=np.zeros((n, E)), gsum=np.zeros(2), dY=np.zeros((self.rows, 32), np.float32))
        self._chk(lib().ftr_learner_statistics(self.engine.h, self.h, N, int(seed), out["c"].ctypes.data,
                                               out["ly"].ctypes.data, out["path"].ctypes.data,
                                               out["gC"].ctypes.data, out["gS"].ctypes.data,
                                               out["gsum"].ctypes.data, out["dY"].ctypes.data))
        return out

    def tables(self):
        n, E = 1 << self.E, self.E
        lw, lratio, lia, lrad, ldiff = (np.zeros((n, E)) for _ in range(5))
        lZ = np.zeros(n)
        self._chk(lib().ftr_learner_tables(self.engine.h, self.h, lw.ctypes.data, lZ.ctypes.data,
                                           lratio.ctypes.data, lia.ctypes.data, lrad.ctypes.data,
                                           ldiff.ctypes.data))
        return dict(lw=lw, lZ=lZ, lratio=lratio, lia=lia, lrad=lrad, ldiff=ldiff)

    def integrate(self, N, seed=0, lanes=0):
        m = ftr_moments()
        opts = ftr_opts(0, -1, int(lanes))
        self._chk(lib().ftr_learner_integrate(self.engine.h, self.h, int(N), int(seed), C.byref(opts),
                                              C.byref(m)))
        return m

    def set_objective(self, objective):
        """0: forward KL (default), 1: variance of the estimator."""
        self._chk(lib().ftr_learner_set_objective(self.h, int(objective)))

    def select(self, n_probe=4_000_000, seed=12345):
        """Plain sampling judges the network: returns (use_learned, eff_plain,
        eff_learned); integrate() then uses the better proposal."""
        u, ep, el = C.c_int(), C.c_double(), C.c_double()
        self._chk(lib().ftr_learner_select(self.engine.h, self.h, int(n_probe), int(seed), C.byref(u),
                                           C.byref(ep), C.byref(el)))
        return bool(u.value), ep.value, el.value

    def sample(self, first, n, seed, X_ptr, w_ptr, path_ptr):
        """Device pointers of n*E doubles, n doubles and n uint64."""
        self._chk(lib().ftr_learner_sample(self.engine.h, self.h, int(first), int(n), int(seed),
                                           C.c_void_p(X_ptr), C.c_void_p(w_ptr), C.c_void_p(path_ptr)))


def integral(engine, prob):
    """The reference's result list [[(re, dre), (im, dim)], ...] for a problem dict."""
    p = engine.prepare(prob)
    m = engine.integrate(p, prob["N"], prob["seed"])
    r = p.finalize(m)
    return [[(r[k, 0, 0], r[k, 0, 1]), (r[k, 1, 0], r[k, 1, 1])] for k in range(p.K)], p.info.IGtr


def to_json(prob):
    return json.dumps(prob)
