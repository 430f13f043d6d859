"""ctypes access to the CPU oracle (oracle/libftr_oracle*.so) and, where it
was built in this container, to the compiled reference harness
(oracle/_ref/libref_harness.so).  Test infrastructure only."""

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def problem_arrays(prob):
    edges = np.array([[e[0][0], e[0][1]] for e in prob["graph"]], dtype=np.int32).reshape(-1)
    nu = np.array([float(e[1]) for e in prob["graph"]], dtype=np.float64)
    P = np.array(prob["scalarproducts"], dtype=np.float64)
    m2 = np.array(prob["masses_sqr"], dtype=np.float64)
    return edges, nu, P, m2


class _Lib:
    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.p = prefix
        f = self.fn
        f("create").restype = C.c_void_p
        f("create").argtypes = [C.c_int, _i32p, _f64p, C.c_int, _f64p, _f64p, C.c_double,
                                C.c_int, C.c_double, C.c_char_p, C.c_int]
        f("destroy").argtypes = [C.c_void_p]
        f("info").argtypes = [C.c_void_p, _f64p]
        f("jtable").argtypes = [C.c_void_p, _f64p, _f64p, _u8p, _u8p]
        f("sample").argtypes = [C.c_void_p, _u64p, C.c_uint64, _f64p, _f64p]

    def fn(self, name):
        return getattr(self.lib, self.p + name)


class Problem:
    """One problem instance inside the oracle (or the reference harness)."""

    def __init__(self, lib, prob):
        self.L = lib
        self.prob = prob
        edges, nu, P, m2 = problem_arrays(prob)
        err = C.create_string_buffer(4096)
        self.E = len(nu)
        self.h = lib.fn("create")(self.E, edges, nu, P.shape[0], np.ascontiguousarray(P.reshape(-1)),
                                  m2, float(prob["dimension"]), int(prob["num_eps_terms"]),
                                  float(prob["lambda"]), err, 4096)
        if not self.h:
            raise ValueError(err.value.decode())
        self.K = int(prob["num_eps_terms"])
        info = np.zeros(13)
        lib.fn("info")(self.h, info)
        self.info = info
        self.V = int(info[1])
        self.L_loops = int(info[2])
        self.W = info[3]
        self.IGtr = info[4]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.fn("destroy")(self.h)
            self.h = None

    def jtable(self):
        n = 1 << self.E
        Jr, r = np.zeros(n), np.zeros(n)
        mm, Lt = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        self.L.fn("jtable")(self.h, Jr, r, mm, Lt)
        return Jr, r, mm, Lt

    def sample(self, state, n):
        st = np.array(state, dtype=np.uint64)
        X = np.zeros(n * self.E)
        px = np.zeros(2 * n)
        self.L.fn("sample")(self.h, st, n, X, px)
        return X.reshape(n, self.E), px.reshape(n, 2), st


class Oracle(_Lib):
    """variant: "strict" (IEEE double, no contraction), "fma" (contracted
    double) or "ld" (x87 long double: the ground truth of the mask)."""

    def __init__(self, variant="strict", fma=False):
        if fma:
            variant = "fma"
        name = {"strict": "libftr_oracle.so", "fma": "libftr_oracle_fma.so",
                "ld": "libftr_oracle_ld.so"}[variant]
        super().__init__(os.path.join(ORACLE_DIR, name), "or_")
        f = self.fn
        f("replay").argtypes = [C.c_void_p, C.c_uint64, _f64p, _f64p, _f64p, _u8p, _i32p]
        f("estimate").restype = C.c_uint64
        f("estimate").argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, _f64p]
        f("finalize").argtypes = [C.c_void_p, _f64p, _f64p]
        f("xoshiro_seed").argtypes = [C.c_uint64, _u64p]
        f("xoshiro_next").argtypes = [_u64p, C.c_uint64, C.c_void_p]
        f("xoshiro_advance").argtypes = [_u64p, C.c_uint64, C.c_uint64]
        f("xoshiro_charpoly").argtypes = [_u64p]
        f("xoshiro_jumppoly_pow2").argtypes = [C.c_int, _u64p]
        f("xoshiro_jump_pow2").argtypes = [C.c_int, _u64p]
        f("fp_sum").restype = C.c_double
        f("fp_sum").argtypes = [_f64p, C.c_uint64]

    def problem(self, prob):
        return OracleProblem(self, prob)

    def seed(self, seed):
        st = np.zeros(4, np.uint64)
        self.fn("xoshiro_seed")(seed, st)
        return st

    def next(self, st, n):
        out = np.zeros(n, np.uint64)
        self.fn("xoshiro_next")(st, n, out.ctypes.data_as(C.c_void_p))
        return out

    def advance(self, st, d):
        self.fn("xoshiro_advance")(st, d & (2**64 - 1), d >> 64)
        return st


class OracleProblem(Problem):
    def replay(self, X, px):
        n = X.shape[0]
        out = np.zeros(n * self.K * 2)
        drop = np.zeros(n, np.uint8)
        vs = np.zeros(n, np.int32)
        self.L.fn("replay")(self.h, n, np.ascontiguousarray(X.reshape(-1)),
                            np.ascontiguousarray(px.reshape(-1)), out, drop, vs)
        return out.reshape(n, self.K, 2), drop, vs

    def estimate(self, N, seed=0, chunks=1, threads=1):
        m = np.zeros(self.K * 6)
        d = self.L.fn("estimate")(self.h, N, seed, chunks, threads, m)
        return m.reshape(self.K, 2, 3), int(d)

    def finalize(self, moments):
        res = np.zeros(self.K * 4)
        self.L.fn("finalize")(self.h, np.ascontiguousarray(moments.reshape(-1)), res)
        return res.reshape(self.K, 2, 2)


REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_harness.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "feyntrop_ref")


def ref_available():
    return os.path.exists(REF_LIB)


class Ref(_Lib):
    def __init__(self):
        super().__init__(REF_LIB, "ref_")
        f = self.fn
        f("values").argtypes = [C.c_void_p, _u64p, C.c_uint64, _f64p, _u8p]
        f("estimate").argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _f64p]

    def problem(self, prob):
        return RefProblem(self, prob)


class RefProblem(Problem):
    def values(self, state, n):
        st = np.array(state, dtype=np.uint64)
        out = np.zeros(n * self.K * 2)
        drop = np.zeros(n, np.uint8)
        self.L.fn("values")(self.h, st, n, out, drop)
        return out.reshape(n, self.K, 2), drop, st

    def estimate(self, N, seed=0, threads=1):
        m = np.zeros(self.K * 6)
        self.L.fn("estimate")(self.h, N, seed, threads, m)
        return m.reshape(self.K, 2, 3)
