"""FP64 CPU oracle of the plmDCA path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle_plmdca.so (plmdca_oracle.c, a plain-C restatement of
SuperDCA's per-locus solve; citations in that file).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may use it, and only as the checker or the timed CPU
baseline; the product (superdca_amd, libsdca.so) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_plmdca.so")


class Stats(C.Structure):
    _fields_ = [("fval", C.c_double), ("gnorm", C.c_double), ("nfeval", C.c_int32), ("iter", C.c_int32),
                ("status", C.c_int32), ("pad", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
        L.orc_nparams.argtypes = [i32, i32]
        L.orc_nparams.restype = C.c_size_t
        L.orc_ref_index.argtypes = [i32, i32, i32, i32, i32]
        L.orc_ref_index.restype = C.c_size_t
        L.orc_hamming.argtypes = [vp, i32, i32, vp, i32]
        L.orc_weights.argtypes = [vp, i32, i32, dbl, vp, vp, vp, i32]
        L.orc_lambda.argtypes = [vp, i32, dbl, dbl, vp, vp, vp]
        L.orc_objective.argtypes = [vp, i32, i32, i32, i32, vp, dbl, dbl, vp, vp]
        L.orc_objective.restype = dbl
        L.orc_blocks_build.argtypes = [vp, i32, i32, i32]
        L.orc_blocks_build.restype = vp
        L.orc_blocks_free.argtypes = [vp]
        L.orc_blocks_npatterns.argtypes = [vp]
        L.orc_blocks_npatterns.restype = C.c_int64
        L.orc_objective_blocked.argtypes = [vp, i32, vp, dbl, dbl, vp, vp, vp]
        L.orc_objective_blocked.restype = dbl
        L.orc_score_row.argtypes = [vp, i32, i32, i32, i32, vp]
        L.orc_ising_gauge.argtypes = [vp, i32, i32, vp]
        L.orc_frobenius_excl.argtypes = [vp, i32, i32]
        L.orc_frobenius_excl.restype = dbl
        L.orc_nom_block.argtypes = [vp, i32, i32, i32, i32, vp]
        L.orc_ising_gauge_float.argtypes = [vp, i32, vp]
        L.orc_frobenius_excl_float.argtypes = [vp, i32, i32]
        L.orc_frobenius_excl_float.restype = C.c_float
        L.orc_nom_pair.argtypes = [vp, vp, i32, i32, vp]
        L.orc_solve_loci.argtypes = [vp, i32, i32, i32, i32, vp, dbl, dbl, vp, i32, dbl, i32, i32, i32, i32,
                                     vp, vp, vp, vp]
        L.orc_solve_loci.restype = i32
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def nparams(L: int, q: int) -> int:
    return int(lib().orc_nparams(L, q))


def ref_index(L: int, q: int, j: int, a: int, s: int) -> int:
    return int(lib().orc_ref_index(L, q, j, a, s))


def hamming(states: np.ndarray, nthreads: int = 4) -> np.ndarray:
    st = np.ascontiguousarray(states, dtype=np.uint8)
    B, L = st.shape
    d = np.zeros(max(1, B * (B - 1) // 2), dtype=np.uint32)
    lib().orc_hamming(_p(st), B, L, _p(d), nthreads)
    return d[: B * (B - 1) // 2]


def weights(states: np.ndarray, theta: float = 0.9, mult=None, nthreads: int = 4):
    st = np.ascontiguousarray(states, dtype=np.uint8)
    B, L = st.shape
    w = np.zeros(B)
    iw = np.zeros(B, dtype=np.uint64)
    m = None if mult is None else np.ascontiguousarray(mult, dtype=np.uint32)
    lib().orc_weights(_p(st), B, L, float(theta), _p(m), _p(w), _p(iw), nthreads)
    return w, iw


def lambdas(w: np.ndarray, lambda_h: float = -1.0, lambda_J: float = -1.0):
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = [C.c_double(), C.c_double(), C.c_double()]
    lib().orc_lambda(_p(w), w.size, float(lambda_h), float(lambda_J), *(C.byref(o) for o in out))
    return tuple(o.value for o in out)  # B_eff, lambda_h, lambda_J


def objective(states, r, w, lh, lJ, x, q=4):
    st = np.ascontiguousarray(states, dtype=np.uint8)
    B, L = st.shape
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(nparams(L, q))
    f = lib().orc_objective(_p(st), B, L, q, int(r), _p(w), float(lh), float(lJ), _p(x), _p(g))
    return f, g


class Blocks:
    """Block-compressed alignment (unique 16-locus patterns), plmDCA_objective.hpp:89-121."""

    def __init__(self, states, q=4):
        self.states = np.ascontiguousarray(states, dtype=np.uint8)
        self.B, self.L = self.states.shape
        self.q = q
        self.h = lib().orc_blocks_build(_p(self.states), self.B, self.L, q)
        self.work = np.zeros(2 * self.B * q)

    def __del__(self):
        try:
            lib().orc_blocks_free(self.h)
        except Exception:
            pass

    @property
    def npatterns(self) -> int:
        return int(lib().orc_blocks_npatterns(self.h))

    def objective(self, r, w, lh, lJ, x):
        w = np.ascontiguousarray(w, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(nparams(self.L, self.q))
        f = lib().orc_objective_blocked(self.h, int(r), _p(w), float(lh), float(lJ), _p(x), _p(g), _p(self.work))
        return f, g


def score_row(x, L, q, r, gap=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    row = np.zeros(L, dtype=np.float32)
    lib().orc_score_row(_p(x), L, q, int(r), q - 1 if gap is None else gap, _p(row))
    return row


def ising_gauge(M, transpose=False):
    M = np.ascontiguousarray(M, dtype=np.float64)
    q = M.shape[0]
    out = np.zeros((q, q))
    lib().orc_ising_gauge(_p(M), q, int(transpose), _p(out))
    return out


def nom_block(x, L, q, r, n):
    x = np.ascontiguousarray(x, dtype=np.float64)
    M = np.zeros((q, q), dtype=np.float32)
    lib().orc_nom_block(_p(x), int(L), int(q), int(r), int(n), _p(M))
    return M


def ising_gauge_float(M):
    M = np.ascontiguousarray(M, dtype=np.float32)
    out = np.zeros_like(M)
    lib().orc_ising_gauge_float(_p(M), M.shape[0], _p(out))
    return out


def nom_pair(Mrn, Mnr, gap=None):
    a = np.ascontiguousarray(Mrn, dtype=np.float32)
    b = np.ascontiguousarray(Mnr, dtype=np.float32)
    q = a.shape[0]
    out = np.zeros(3, dtype=np.float32)
    lib().orc_nom_pair(_p(a), _p(b), q, q - 1 if gap is None else int(gap), _p(out))
    return out


def frobenius_excl(G, exclude):
    G = np.ascontiguousarray(G, dtype=np.float64)
    return float(lib().orc_frobenius_excl(_p(G), G.shape[0], int(exclude)))


def solve_loci(states, w, lh, lJ, loci, q=4, gap=None, gtol=1e-3, max_iter=2000, m=10, nthreads=4, blocked=True,
               x0=None, want_rows=True, want_x=False):
    """Threaded FP64 solve of the listed loci -> (rows, x, stats dict)."""
    st = np.ascontiguousarray(states, dtype=np.uint8)
    B, L = st.shape
    loci = np.ascontiguousarray(loci, dtype=np.int32)
    K = loci.size
    n = nparams(L, q)
    rows = np.zeros((K, L), dtype=np.float32) if want_rows else None
    X = np.zeros((K, n)) if want_x else None
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).reshape(K, n)
    sts = (Stats * max(K, 1))()
    lib().orc_solve_loci(_p(st), B, L, q, q - 1 if gap is None else gap, _p(np.ascontiguousarray(w, dtype=np.float64)),
                         float(lh), float(lJ), _p(loci), K, float(gtol), int(max_iter), int(m), int(nthreads),
                         int(blocked), _p(x0a), _p(rows), _p(X), C.cast(sts, C.c_void_p))
    stats = {k: np.array([getattr(sts[i], k) for i in range(K)]) for k in ("fval", "gnorm", "nfeval", "iter", "status")}
    return rows, X, stats
