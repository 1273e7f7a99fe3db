"""ctypes binding of the CPU oracle (oracle/_build/libbk_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it, and only as the checker or the timed CPU baseline.
The oracle exposes the same learner interface as the GPU library (bko_*
instead of bk_*), over the same C structs.
"""
import ctypes as C
import os

import numpy as np

import bankai_gpu as bg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libbk_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        raise RuntimeError(f"{ORACLE_LIB} is missing: run `make -C oracle`")
    lib = C.CDLL(ORACLE_LIB)
    P = C.c_void_p
    sig = {
        "bko_last_error": ([], C.c_char_p),
        "bko_learner_create": ([C.POINTER(bg.DatasetView), C.POINTER(bg.LearnerParamsC), C.c_int, C.POINTER(P)],
                               C.c_int),
        "bko_learner_train": ([P, P, P, P, C.c_int64, C.POINTER(bg.TreeOutC)], C.c_int),
        "bko_learner_add_prediction_to_score": ([P, C.POINTER(bg.TreeOutC), P], C.c_int),
        "bko_learner_add_out_of_bag_prediction": ([P, C.POINTER(bg.TreeOutC), P, C.c_int64, P], C.c_int),
        "bko_learner_leaf_rows": ([P, C.c_int, P, C.POINTER(C.c_int64)], C.c_int),
        "bko_learner_stats": ([P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "bko_learner_quant_state": ([P, P, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        "bko_learner_free": ([P], None),
        "bko_learner_root_sums": ([P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
        "bko_block_info": ([C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], None),
        "bko_learner_histogram_cache_capacity": ([P], C.c_int64),
        "bko_learner_reset_histogram_pool_config": ([P], None),
        "bko_learner_continue_after_dataset_reset": ([P, P], None),
        "bko_mt_new": ([C.c_uint32], P),
        "bko_mt_next": ([P], C.c_uint32),
        "bko_mt_uniform_real": ([P], C.c_double),
        "bko_mt_uniform_int": ([P, C.c_uint32], C.c_uint32),
        "bko_mt_free": ([P], None),
        "bko_discretize": ([P, P, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, P,
                            C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        "bko_noise": ([C.c_int64, C.c_uint32, C.c_int, P, P], None),
        "bko_hist_f64": ([P, P, C.c_int64, P, P, C.c_int, P], None),
        "bko_search": ([C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, C.c_double, C.c_double,
                        C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_double, P, P, P, P], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleLearner:
    """Same interface as bankai_gpu.Learner, computed by the CPU restatement."""

    def __init__(self, dataset, params=None, is_constant_hessian=False):
        self._lib = load()
        self.params = params or bg.LearnerParams()
        self.dataset = dataset
        p = self.params.to_c()
        out = C.c_void_p()
        rc = self._lib.bko_learner_create(C.byref(dataset.view), C.byref(p), int(is_constant_hessian), C.byref(out))
        if rc != 0:
            raise RuntimeError(self._lib.bko_last_error().decode())
        self.handle = out
        self.tree = bg.TreeBuffers(self.params.num_leaves)

    def train(self, gradients, hessians, rows=None):
        g = np.ascontiguousarray(gradients, dtype=np.float32)
        h = np.ascontiguousarray(hessians, dtype=np.float32)
        r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
        rc = self._lib.bko_learner_train(self.handle, _p(g), _p(h), _p(r), 0 if r is None else len(r),
                                         C.byref(self.tree.c))
        if rc != 0:
            raise RuntimeError(self._lib.bko_last_error().decode())
        return self.tree.as_dict()

    def add_prediction_to_score(self, leaf_values, score):
        self.tree.set_leaf_values(leaf_values)
        self._lib.bko_learner_add_prediction_to_score(self.handle, C.byref(self.tree.c), _p(score))

    def add_out_of_bag_prediction(self, leaf_values, rows, score):
        self.tree.set_leaf_values(leaf_values)
        r = np.ascontiguousarray(rows, dtype=np.int32)
        self._lib.bko_learner_add_out_of_bag_prediction(self.handle, C.byref(self.tree.c), _p(r), len(r), _p(score))

    def leaf_rows(self, leaf):
        out = np.zeros(self.dataset.num_data, np.int32)
        n = C.c_int64()
        self._lib.bko_learner_leaf_rows(self.handle, leaf, _p(out), C.byref(n))
        return out[: n.value].copy()

    def stats(self):
        a, b = C.c_int64(), C.c_int64()
        self._lib.bko_learner_stats(self.handle, C.byref(a), C.byref(b))
        return {"hist_rows": a.value, "part_rows": b.value}

    def quant_state(self):
        packed = np.zeros(self.dataset.num_data, np.int16)
        gs, hs = C.c_double(), C.c_double()
        self._lib.bko_learner_quant_state(self.handle, _p(packed), C.byref(gs), C.byref(hs))
        return packed, gs.value, hs.value

    def root_sums(self):
        """(sum_g, sum_h, packed int sum) of the last tree's root (before_train)."""
        sg, sh, isum = C.c_double(), C.c_double(), C.c_int64()
        self._lib.bko_learner_root_sums(self.handle, C.byref(sg), C.byref(sh), C.byref(isum))
        return sg.value, sh.value, isum.value

    def histogram_cache_capacity(self):
        return self._lib.bko_learner_histogram_cache_capacity(self.handle)

    def reset_histogram_pool_config(self):
        self._lib.bko_learner_reset_histogram_pool_config(self.handle)

    def continue_after_dataset_reset(self, previous):
        self._lib.bko_learner_continue_after_dataset_reset(self.handle, previous.handle)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.bko_learner_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Mt19937:
    def __init__(self, seed=5489):
        self._lib = load()
        self.h = self._lib.bko_mt_new(seed)

    def next(self):
        return self._lib.bko_mt_next(self.h)

    def uniform_real(self):
        return self._lib.bko_mt_uniform_real(self.h)

    def uniform_int(self, n):
        return self._lib.bko_mt_uniform_int(self.h, n)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.bko_mt_free(self.h)


def discretize(g, h, bins, const_hess=False, stochastic=False, seed=0, num_threads=1, draws=0):
    lib = load()
    g = np.ascontiguousarray(g, np.float32)
    h = np.ascontiguousarray(h, np.float32)
    packed = np.zeros(len(g), np.int16)
    gs, hs = C.c_double(), C.c_double()
    rc = lib.bko_discretize(_p(g), _p(h), len(g), bins, int(const_hess), int(stochastic), seed, num_threads, draws,
                            _p(packed), C.byref(gs), C.byref(hs))
    if rc != 0:
        return None
    return packed, gs.value, hs.value


def block_info(num_threads, cnt, min_cnt):
    """Threading::BlockInfo as the oracle's noise layout computes it: (blocks, block size)."""
    nb, bs = C.c_int64(), C.c_int64()
    load().bko_block_info(int(num_threads), int(cnt), int(min_cnt), C.byref(nb), C.byref(bs))
    return nb.value, bs.value


def noise(n, seed, num_threads):
    lib = load()
    gn = np.zeros(n)
    hn = np.zeros(n)
    lib.bko_noise(n, seed, num_threads, _p(gn), _p(hn))
    return gn, hn


def hist_f64(bins, rows, og, oh, const_hess, num_bin):
    lib = load()
    out = np.zeros(2 * num_bin)
    rows_a = None if rows is None else np.ascontiguousarray(rows, np.int32)
    n = len(og)
    lib.bko_hist_f64(_p(np.ascontiguousarray(bins, np.uint8)), _p(rows_a), n, _p(np.ascontiguousarray(og, np.float32)),
                     _p(np.ascontiguousarray(oh, np.float32)), int(const_hess), _p(out))
    return out


def search(hist, num_bin, missing_type=0, default_bin=0, most_freq_bin=0, bin_type=0, l1=0.0, l2=0.0, mds=0.0,
           smooth=0.0, min_hess=1e-3, min_gain=0.0, cat_smooth=10.0, cat_l2=10.0, min_data=20, max_cat_threshold=32,
           max_cat_to_onehot=4, min_data_per_group=100, sum_g=0.0, sum_h=0.0, int_g=0, int_h=0, gs=1.0, hs=1.0,
           num_data=0, parent=0.0):
    """One threshold search; `hist` is an (n, 2) float64 (f64 search) or int64 (integer search) array."""
    lib = load()
    mode = 1 if hist.dtype == np.int64 else 0
    hist = np.ascontiguousarray(hist)
    dp = np.array([l1, l2, mds, smooth, min_hess, min_gain, cat_smooth, cat_l2])
    ip = np.array([min_data, max_cat_threshold, max_cat_to_onehot, min_data_per_group], np.int32)
    iout = np.zeros(5, np.int32)
    dout = np.zeros(7)
    cat = np.zeros(256, np.uint32)
    ncat = np.zeros(1, np.int32)
    lib.bko_search(mode, _p(hist), num_bin, missing_type, default_bin, most_freq_bin, bin_type, _p(dp), _p(ip),
                   sum_g, sum_h, int_g, int_h, gs, hs, num_data, parent, _p(iout), _p(dout), _p(cat), _p(ncat))
    return {"threshold": int(iout[0]), "default_left": bool(iout[1]), "left_count": int(iout[2]),
            "right_count": int(iout[3]), "splittable": bool(iout[4]), "gain": dout[0], "left_output": dout[1],
            "right_output": dout[2], "left_sum_gradient": dout[3], "left_sum_hessian": dout[4],
            "right_sum_gradient": dout[5], "right_sum_hessian": dout[6], "cat": cat[: ncat[0]].copy()}


def root_hist_f64(dataset, g, h, const_hess):
    """The reference's root histogram in FP64 (histogram.rs:117-245 per group, rows in
    order, serial.rs:982-988 for a constant hessian): (total_bins, 2) float64."""
    cols = dataset.group_columns()
    nbins = dataset.group_num_total_bin()
    parts = []
    for col, nb in zip(cols, nbins):
        out = hist_f64(col, None, g, h, const_hess, int(nb)).reshape(-1, 2)
        if const_hess:
            out[:, 1] *= float(np.float32(h[0]))
        parts.append(out)
    return np.concatenate(parts)
