"""superdca_amd -- MI355X plmDCA engine (Python face of libsdca.so).

The compute path is libsdca.so (HIP kernels for gfx950 behind the C ABI declared in
include/sdca.h).  This module binds that ABI with ctypes and mirrors the reference's
per-locus interface (SuperDCA/include/plmDCA.hpp:436 run_plmDCA, the TBB body
plmDCA_solver at :199-425 and the objective of plmDCA_objective.hpp:38-249) so that
tests and benchmarks call exactly what a C++ caller would.  There is no fallback: if
the library is missing or cannot be loaded every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdca.so")

SDCA_OK = 0
SDCA_E_ARG = -1
SDCA_E_HIP = -2
SDCA_E_NONFINITE = -3
SDCA_E_NOCONV = -4
SDCA_E_STATE = -5
SDCA_E_IO = -6

PREC_FP32 = 0
PREC_SPLIT_BF16 = 1
# sdca_opts.formulation: AUTO picks GRAM when there is no warm start and B <= K_x
FORM_AUTO = 0
FORM_PRIMAL = 1
FORM_GRAM = 2

ST_CONVERGED = 0
ST_ITER_CAP = 1
ST_LS_FAIL = 2
ST_NONFINITE = 3

WARM_DIRECTION = 0
WARM_POINT = 1
WARM_NO_FIELDS = 2

# Every symbol include/sdca.h declares.
EXPORTED = (
    "sdca_default_opts", "sdca_last_error", "sdca_ctx_create", "sdca_ctx_destroy", "sdca_num_columns",
    "sdca_reweight", "sdca_set_weights", "sdca_eval_batch", "sdca_solve_loci", "sdca_get_timing",
    "sdca_prepare_alignment", "sdca_write_couplings", "sdca_pair_counts", "sdca_pair_count_table",
    "sdca_mi_topk", "sdca_pair_features", "sdca_mlp_init", "sdca_mlp_num_params", "sdca_mlp_get_params",
    "sdca_mlp_set_params", "sdca_mlp_forward", "sdca_mlp_loss_grad", "sdca_mlp_train", "sdca_solve_loci_mlp",
    "sdca_gemm_config", "sdca_write_couplings_top", "sdca_set_scan_split", "sdca_write_scan_couplings",
    "sdca_reweight_noalloc", "sdca_store_create", "sdca_store_destroy", "sdca_store_put_rows", "sdca_store_put_blocks",
    "sdca_store_write_couplings", "sdca_store_write_norm_of_mean", "sdca_store_write_norm_of_mean_scan",
    "sdca_multi_create", "sdca_multi_destroy", "sdca_multi_solve", "sdca_multi_get_timing",
)


class SdcaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sdca error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("precision", C.c_int32), ("batch_loci", C.c_int32),
                ("history", C.c_int32), ("max_iter", C.c_int32), ("max_ls", C.c_int32),
                ("grad_threshold", C.c_double), ("verbose", C.c_int32), ("autotune", C.c_int32),
                ("gemm_ksplit", C.c_int32), ("gemm_gm_energy", C.c_int32), ("gemm_gm_grad", C.c_int32),
                ("gemm_tn_energy", C.c_int32), ("gemm_tn_grad", C.c_int32), ("pipeline_groups", C.c_int32),
                ("coarse_factor", C.c_int32), ("fine_factor", C.c_int32), ("j_threshold", C.c_double),
                ("formulation", C.c_int32), ("graph_rounds", C.c_int32), ("subspace_rank", C.c_int32),
                ("subspace_threshold", C.c_double)]


class LocusStats(C.Structure):
    _fields_ = [("fval", C.c_double), ("gnorm", C.c_double), ("nfeval", C.c_int32), ("iter", C.c_int32),
                ("status", C.c_int32), ("nfeval_first", C.c_int32), ("gnorm_J", C.c_double),
                ("nfeval_fine", C.c_int32), ("nfeval_sub", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("gemm_energy_ms", C.c_double), ("gemm_grad_ms", C.c_double), ("softmax_ms", C.c_double),
                ("lbfgs_ms", C.c_double), ("score_ms", C.c_double), ("reweight_ms", C.c_double),
                ("total_ms", C.c_double), ("n_evals", C.c_int64), ("n_energy_launches", C.c_int64),
                ("n_grad_launches", C.c_int64), ("flops_energy", C.c_double), ("flops_grad", C.c_double),
                ("bytes_lbfgs", C.c_double), ("update_ms", C.c_double), ("dots_ms", C.c_double),
                ("bytes_update", C.c_double), ("bytes_dots", C.c_double), ("subspace_ms", C.c_double),
                ("n_sub_evals", C.c_int64), ("flops_sub", C.c_double), ("host_ms", C.c_double),
                ("gstep_ms", C.c_double), ("bytes_gstep", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """Loads libsdca.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SdcaError(SDCA_E_STATE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.sdca_last_error.restype = C.c_char_p
    L.sdca_default_opts.argtypes = [P(Opts)]
    L.sdca_ctx_create.argtypes = [P(C.c_void_p), C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(Opts)]
    L.sdca_ctx_destroy.argtypes = [C.c_void_p]
    L.sdca_num_columns.argtypes = [C.c_void_p]
    L.sdca_num_columns.restype = C.c_int64
    L.sdca_reweight.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sdca_set_weights.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
    L.sdca_eval_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sdca_solve_loci.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sdca_get_timing.argtypes = [C.c_void_p, P(Timing), C.c_int32]
    L.sdca_prepare_alignment.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                         C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sdca_write_couplings.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    L.sdca_write_couplings_top.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int64,
                                           C.c_double]
    L.sdca_set_scan_split.argtypes = [C.c_void_p, C.c_int32]
    L.sdca_write_scan_couplings.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                            C.c_int32]
    L.sdca_pair_counts.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.sdca_pair_count_table.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
    L.sdca_mi_topk.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    L.sdca_pair_features.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.sdca_mlp_init.argtypes = [C.c_void_p, C.c_int32, C.c_uint64]
    L.sdca_mlp_num_params.argtypes = [C.c_void_p]
    L.sdca_mlp_num_params.restype = C.c_int64
    L.sdca_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.sdca_mlp_set_params.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]
    L.sdca_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.sdca_mlp_loss_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_void_p]
    L.sdca_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                 C.c_float, C.c_uint64, C.c_void_p]
    L.sdca_solve_loci_mlp.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    L.sdca_gemm_config.argtypes = [C.c_void_p, C.c_void_p]
    L.sdca_store_create.argtypes = [P(C.c_void_p), C.c_char_p, C.c_int64, C.c_int64, C.c_int32]
    L.sdca_store_destroy.argtypes = [C.c_void_p]
    L.sdca_store_put_rows.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
    L.sdca_store_put_blocks.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    L.sdca_store_write_couplings.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_int32, C.c_int64,
                                             C.c_double, C.c_int32]
    L.sdca_store_write_norm_of_mean.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                C.c_int32]
    L.sdca_store_write_norm_of_mean_scan.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int32,
                                                     C.c_int32, C.c_int32]
    L.sdca_multi_create.argtypes = [P(C.c_void_p), C.c_void_p, C.c_int32, P(Opts)]
    L.sdca_multi_destroy.argtypes = [C.c_void_p]
    L.sdca_multi_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sdca_multi_get_timing.argtypes = [C.c_void_p, P(Timing), C.c_int32]
    L.sdca_reweight_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
    for name in ("sdca_ctx_create", "sdca_reweight", "sdca_set_weights", "sdca_eval_batch", "sdca_solve_loci",
                 "sdca_get_timing", "sdca_prepare_alignment", "sdca_write_couplings", "sdca_pair_counts",
                 "sdca_pair_count_table", "sdca_mi_topk", "sdca_pair_features", "sdca_mlp_init",
                 "sdca_mlp_get_params", "sdca_mlp_set_params", "sdca_mlp_forward", "sdca_mlp_loss_grad",
                 "sdca_mlp_train", "sdca_solve_loci_mlp", "sdca_gemm_config", "sdca_write_couplings_top",
                 "sdca_set_scan_split", "sdca_write_scan_couplings", "sdca_reweight_noalloc", "sdca_store_create",
                 "sdca_store_put_rows", "sdca_store_put_blocks", "sdca_store_write_couplings",
                 "sdca_store_write_norm_of_mean", "sdca_store_write_norm_of_mean_scan", "sdca_multi_create",
                 "sdca_multi_solve", "sdca_multi_get_timing"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc: int, allow=()) -> int:
    if rc != SDCA_OK and rc not in allow:
        raise SdcaError(rc, lib().sdca_last_error().decode(errors="replace"))
    return rc


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().sdca_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


def nparams(L: int, q: int) -> int:
    """Length of one locus' parameter vector in the reference layout (L*q*q + q)."""
    return L * q * q + q


class Context:
    """One alignment on one GPU: the role of run_plmDCA's inputs (plmDCA.hpp:436-437)."""

    def __init__(self, states: np.ndarray, q: int = 4, gap: int | None = None, **opts):
        states = np.ascontiguousarray(states, dtype=np.uint8)
        if states.ndim != 2:
            raise ValueError("states must be B x L")
        self.B, self.L = states.shape
        self.q = q
        self.gap = q - 1 if gap is None else gap
        self.opts = default_opts(**opts)
        h = C.c_void_p()
        _check(lib().sdca_ctx_create(C.byref(h), _ptr(states), self.B, self.L, q, self.gap, C.byref(self.opts)))
        self._h = h
        self.B_eff = self.lambda_h = self.lambda_J = None

    def close(self):
        if getattr(self, "_h", None):
            lib().sdca_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def num_columns(self) -> int:
        return int(lib().sdca_num_columns(self._h))

    def set_scan_split(self, n_first: int) -> None:
        """Two-alignment scan (SuperDCA.cpp:155-159): loci [0, n_first) are the first
        alignment's and are fitted on the remaining columns only; 0 turns it off."""
        _check(lib().sdca_set_scan_split(self._h, int(n_first)))

    def reweight(self, theta: float = 0.9, mult: np.ndarray | None = None):
        """calculate_weights (plmDCA_optimizer_parameters.hpp:39-98) -> (w, int_w)."""
        w = np.zeros(self.B, dtype=np.float64)
        iw = np.zeros(self.B, dtype=np.uint64)
        m = None if mult is None else np.ascontiguousarray(mult, dtype=np.uint32)
        _check(lib().sdca_reweight(self._h, float(theta), _ptr(m), _ptr(w), _ptr(iw)))
        return w, iw

    def set_weights(self, w: np.ndarray, lambda_h: float = -1.0, lambda_J: float = -1.0):
        """Installs the weights; B_eff and lambda as plmDCA_optimizer_parameters.hpp:257-284
        (negative = automatic) -> (B_eff, lambda_h, lambda_J)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros(3, dtype=np.float64)
        _check(lib().sdca_set_weights(self._h, _ptr(w), float(lambda_h), float(lambda_J), _ptr(out)))
        self.B_eff, self.lambda_h, self.lambda_J = (float(v) for v in out)
        return self.B_eff, self.lambda_h, self.lambda_J

    def eval_batch(self, loci, x: np.ndarray):
        """f and g of plmDCA_objective_fval_and_gradient (plmDCA_objective.hpp:38-249) for
        each locus, x and g in the reference layout."""
        loci = np.ascontiguousarray(loci, dtype=np.int32)
        K = loci.size
        n = nparams(self.L, self.q)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(K, n)
        f = np.zeros(K, dtype=np.float64)
        g = np.zeros((K, n), dtype=np.float64)
        _check(lib().sdca_eval_batch(self._h, _ptr(loci), K, _ptr(x), _ptr(f), _ptr(g)))
        return f, g

    def solve(self, loci, warm: np.ndarray | None = None, want_x: bool = False, rows: bool = True):
        """The per-locus solves of plmDCA_solver::operator() (plmDCA.hpp:277-396) as one
        batched L-BFGS -> (score_rows K x L float32, stats, x or None, rc)."""
        loci = np.ascontiguousarray(loci, dtype=np.int32)
        K = loci.size
        n = nparams(self.L, self.q)
        S = np.zeros((K, self.L), dtype=np.float32) if rows else None
        X = np.zeros((K, n), dtype=np.float64) if want_x else None
        st = (LocusStats * max(K, 1))()
        wm = None if warm is None else np.ascontiguousarray(warm, dtype=np.float64).reshape(K, n)
        rc = _check(lib().sdca_solve_loci(self._h, _ptr(loci), K, _ptr(wm), _ptr(S), _ptr(X),
                                          C.cast(st, C.c_void_p)), allow=(SDCA_E_NOCONV,))
        stats = {
            "fval": np.array([st[k].fval for k in range(K)]),
            "gnorm": np.array([st[k].gnorm for k in range(K)]),
            "nfeval": np.array([st[k].nfeval for k in range(K)], dtype=np.int64),
            "iter": np.array([st[k].iter for k in range(K)], dtype=np.int64),
            "status": np.array([st[k].status for k in range(K)], dtype=np.int64),
            "nfeval_first": np.array([st[k].nfeval_first for k in range(K)], dtype=np.int64),
            "gnorm_J": np.array([st[k].gnorm_J for k in range(K)]),
            "nfeval_fine": np.array([st[k].nfeval_fine for k in range(K)], dtype=np.int64),
            "nfeval_sub": np.array([st[k].nfeval_sub for k in range(K)], dtype=np.int64),
        }
        return S, stats, X, rc

    def pair_counts(self, pairs) -> np.ndarray:
        """Exact q x q co-occurrence counts of listed locus pairs (popcount path)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        out = np.zeros((pairs.shape[0], self.q, self.q), dtype=np.int32)
        _check(lib().sdca_pair_counts(self._h, _ptr(pairs), pairs.shape[0], _ptr(out)))
        return out

    def pair_count_table(self, i0: int, nloci: int) -> np.ndarray:
        """Rows of the full int8-MFMA count table X8 X8^T ((nloci*QP) x (L*QP))."""
        QP = 4 if self.q <= 4 else 8
        out = np.zeros((nloci * QP, self.L * QP), dtype=np.int32)
        _check(lib().sdca_pair_count_table(self._h, int(i0), int(nloci), _ptr(out)))
        return out

    def mi_topk(self, K: int):
        """(partners L x K int32, MI L x K float32): each locus' K largest-MI partners."""
        idx = np.zeros((self.L, K), dtype=np.int32)
        val = np.zeros((self.L, K), dtype=np.float32)
        _check(lib().sdca_mi_topk(self._h, int(K), _ptr(idx), _ptr(val)))
        return idx, val

    # ---- warm starts

    def pair_features(self, pairs) -> np.ndarray:
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        D = 2 * self.q * self.q + 2 * self.q
        out = np.zeros((pairs.shape[0], D), dtype=np.float32)
        _check(lib().sdca_pair_features(self._h, _ptr(pairs), pairs.shape[0], _ptr(out)))
        return out

    def mlp_init(self, hidden: int = 128, seed: int = 0) -> None:
        _check(lib().sdca_mlp_init(self._h, int(hidden), int(seed)))
        self.mlp_hidden = int(hidden)

    def mlp_params(self) -> np.ndarray:
        n = int(lib().sdca_mlp_num_params(self._h))
        out = np.zeros(max(n, 0), dtype=np.float32)
        _check(lib().sdca_mlp_get_params(self._h, _ptr(out), n))
        return out

    def mlp_set_params(self, params: np.ndarray, hidden: int = 128) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().sdca_mlp_set_params(self._h, int(hidden), _ptr(p), p.size))
        self.mlp_hidden = int(hidden)

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.q * self.q), dtype=np.float32)
        _check(lib().sdca_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def mlp_loss_grad(self, x, t, mask=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        loss = np.zeros(1)
        g = np.zeros(int(lib().sdca_mlp_num_params(self._h)), dtype=np.float32)
        _check(lib().sdca_mlp_loss_grad(self._h, _ptr(x), _ptr(t), _ptr(m), x.shape[0], _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def mlp_train(self, x, t, mask=None, epochs=10, batch=4096, lr=1e-3, seed=0) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        losses = np.zeros(epochs)
        _check(lib().sdca_mlp_train(self._h, _ptr(x), _ptr(t), _ptr(m), x.shape[0], int(epochs), int(batch),
                                    float(lr), int(seed), _ptr(losses)))
        return losses

    def solve_mlp(self, loci, npartners: int, mode: int = WARM_DIRECTION, want_x: bool = False, rows: bool = True):
        """Batched solve from network warm starts -> (rows, stats, x, rc) like solve()."""
        loci = np.ascontiguousarray(loci, dtype=np.int32)
        K = loci.size
        n = nparams(self.L, self.q)
        S = np.zeros((K, self.L), dtype=np.float32) if rows else None
        X = np.zeros((K, n), dtype=np.float64) if want_x else None
        st = (LocusStats * max(K, 1))()
        rc = _check(lib().sdca_solve_loci_mlp(self._h, _ptr(loci), K, int(npartners), int(mode), _ptr(S), _ptr(X),
                                              C.cast(st, C.c_void_p)), allow=(SDCA_E_NOCONV,))
        stats = {k: np.array([getattr(st[i], k) for i in range(K)]) for k in ("fval", "gnorm", "nfeval", "iter",
                                                                               "status", "nfeval_first", "gnorm_J",
                                                                               "nfeval_fine")}
        return S, stats, X, rc

    def timing(self, reset: bool = False) -> dict:
        t = Timing()
        _check(lib().sdca_get_timing(self._h, C.byref(t), int(reset)))
        return {name: getattr(t, name) for name, _ in Timing._fields_}

    def gemm_config(self) -> dict:
        """GEMM configuration of the current workspace (after the first evaluation)."""
        cfg = (C.c_int32 * 7)()
        _check(lib().sdca_gemm_config(self._h, cfg))
        return {"tuned": bool(cfg[0]), "ksplit": cfg[1], "kchunk": cfg[2], "gm_energy": cfg[3], "gm_grad": cfg[4],
                "tn_energy": cfg[5], "tn_grad": cfg[6]}


class Multi:
    """Several alignments (Contexts with their weights installed) solved in one batched
    solve (sdca_multi_*): every round's kernels cover the loci of all alignments in flight."""

    def __init__(self, contexts, **opts):
        self.contexts = list(contexts)  # kept alive: the solve reads their device data
        self.opts = default_opts(**opts)
        arr = (C.c_void_p * len(self.contexts))(*[c._h.value for c in self.contexts])
        h = C.c_void_p()
        _check(lib().sdca_multi_create(C.byref(h), C.cast(arr, C.c_void_p), len(self.contexts), C.byref(self.opts)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sdca_multi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, alns, loci):
        """-> (list of score rows, one per item, each the L of its alignment; stats; rc)."""
        alns = np.ascontiguousarray(alns, dtype=np.int32)
        loci = np.ascontiguousarray(loci, dtype=np.int32)
        K = loci.size
        Ls = np.array([self.contexts[a].L for a in alns], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(Ls)]).astype(np.int64)
        S = np.zeros(int(off[-1]), dtype=np.float32)
        st = (LocusStats * max(K, 1))()
        rc = _check(lib().sdca_multi_solve(self._h, _ptr(alns), _ptr(loci), K, _ptr(S), _ptr(off[:-1].copy()),
                                           C.cast(st, C.c_void_p)), allow=(SDCA_E_NOCONV,))
        rows = [S[off[k]:off[k + 1]] for k in range(K)]
        stats = {k: np.array([getattr(st[i], k) for i in range(K)]) for k in ("fval", "gnorm", "nfeval", "iter",
                                                                               "status", "gnorm_J", "nfeval_fine")}
        return rows, stats, rc

    def timing(self, reset: bool = False) -> dict:
        t = Timing()
        _check(lib().sdca_multi_get_timing(self._h, C.byref(t), int(reset)))
        return {name: getattr(t, name) for name, _ in Timing._fields_}


class Store:
    """File-backed rows x cols entries of `entry` floats (sdca_store_*): the scores, or the
    q x q coupling blocks of --norm-of-mean-scoring, streamed to the writers in tiles."""

    def __init__(self, path: str, rows: int, cols: int, entry: int = 1):
        h = C.c_void_p()
        _check(lib().sdca_store_create(C.byref(h), path.encode(), int(rows), int(cols), int(entry)))
        self._h, self.rows, self.cols, self.entry = h, rows, cols, entry

    def close(self):
        if getattr(self, "_h", None):
            lib().sdca_store_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def put_rows(self, r0: int, rows: np.ndarray) -> None:
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        _check(lib().sdca_store_put_rows(self._h, int(r0), rows.shape[0], _ptr(rows)))

    def put_blocks(self, row: int, x: np.ndarray, L: int, q: int, r: int, col_loci) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        cl = np.ascontiguousarray(col_loci, dtype=np.int32)
        _check(lib().sdca_store_put_blocks(self._h, int(row), _ptr(x), int(L), int(q), int(r), _ptr(cl)))

    def write_couplings(self, path_all, path_top, translation, base=1, top_n=0, min_score=-np.inf, tile=0):
        tr = np.ascontiguousarray(translation, dtype=np.int32)
        _check(lib().sdca_store_write_couplings(self._h, None if path_all is None else str(path_all).encode(),
                                                None if path_top is None else str(path_top).encode(), _ptr(tr),
                                                int(base), int(top_n), float(min_score), int(tile)))

    def write_norm_of_mean(self, path, translation, base=1, q=4, gap=None, tile=0):
        tr = np.ascontiguousarray(translation, dtype=np.int32)
        _check(lib().sdca_store_write_norm_of_mean(self._h, str(path).encode(), _ptr(tr), int(base), int(q),
                                                   q - 1 if gap is None else int(gap), int(tile)))

    def write_norm_of_mean_scan(self, path, translation1, translation2, base=1, q=4, gap=None):
        t1 = np.ascontiguousarray(translation1, dtype=np.int32)
        t2 = np.ascontiguousarray(translation2, dtype=np.int32)
        _check(lib().sdca_store_write_norm_of_mean_scan(self._h, str(path).encode(), _ptr(t1), _ptr(t2), int(base),
                                                        int(q), q - 1 if gap is None else int(gap)))


def reweight_noalloc(d_states: int, B: int, L: int, q: int, theta: float, d_mult: int | None, d_int_w: int | None,
                     d_w: int, stream: int | None = None) -> None:
    """sdca_reweight_noalloc on device buffers given by address (no device allocation; the
    kernels are queued on `stream` and the call returns without waiting)."""
    _check(lib().sdca_reweight_noalloc(C.c_void_p(d_states), int(B), int(L), int(q), float(theta),
                                       C.c_void_p(d_mult) if d_mult else None,
                                       C.c_void_p(d_int_w) if d_int_w else None, C.c_void_p(d_w),
                                       C.c_void_p(stream) if stream else None))


def prepare_alignment(chars: np.ndarray, maf_threshold: float = 0.01, gap_threshold: float = 0.15,
                      no_filter: bool = False, q: int = 4):
    """MAF/gap filter and envelope transform (SuperDCA.cpp:274-277,354) -> (states, loci)."""
    chars = np.ascontiguousarray(chars, dtype=np.uint8)
    B, L_in = chars.shape
    states = np.zeros((B, L_in), dtype=np.uint8)
    loci = np.zeros(L_in, dtype=np.int32)
    L = C.c_int32(0)
    _check(lib().sdca_prepare_alignment(_ptr(chars), B, L_in, float(maf_threshold), float(gap_threshold),
                                        int(no_filter), q, _ptr(states), _ptr(loci), C.byref(L)))
    Lk = L.value
    return np.ascontiguousarray(states.reshape(-1)[: B * Lk].reshape(B, Lk)), loci[:Lk].copy()


def write_couplings(path: str, S: np.ndarray, translation: np.ndarray, base: int = 1) -> None:
    """Symmetrised "score i j" lines of plmDCA.hpp:667-693."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    tr = np.ascontiguousarray(translation, dtype=np.int32)
    _check(lib().sdca_write_couplings(path.encode(), _ptr(S), S.shape[0], _ptr(tr), int(base)))


def write_scan_couplings(path: str, S: np.ndarray, translation1: np.ndarray, translation2: np.ndarray,
                         base: int = 1) -> None:
    """The scan's "score r n" lines (plmDCA.hpp:625-638,672-681): S is K x L2."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    t1 = np.ascontiguousarray(translation1, dtype=np.int32)
    t2 = np.ascontiguousarray(translation2, dtype=np.int32)
    if S.shape != (t1.size, t2.size):
        raise ValueError(f"S is {S.shape}, expected ({t1.size}, {t2.size})")
    _check(lib().sdca_write_scan_couplings(path.encode(), _ptr(S), S.shape[0], S.shape[1], _ptr(t1), _ptr(t2),
                                           int(base)))


def write_couplings_top(path: str, S: np.ndarray, translation: np.ndarray, base: int = 1, top_n: int = 0,
                        min_score: float = float("-inf")) -> None:
    """The largest symmetrised scores, largest first, in the line format of write_couplings
    (no counterpart in SuperDCA, whose .all file, plmDCA.hpp:667-693, lists every pair)."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    tr = np.ascontiguousarray(translation, dtype=np.int32)
    _check(lib().sdca_write_couplings_top(path.encode(), _ptr(S), S.shape[0], _ptr(tr), int(base), int(top_n),
                                          float(min_score)))
