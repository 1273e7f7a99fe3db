"""ctypes binding of libbankai_gpu.so: the host-side mirror of the reference's
tree-learner seat.

`Learner` mirrors the method set `Trainer` uses on `SerialTreeLearner`
(crates/lgbm-core/src/treelearner/serial.rs:279-2043): construction from a
binned dataset and a config, `train(gradients, hessians)` returning a tree,
`add_prediction_to_score`, the rows of a leaf, and `reset_constant_hessian`.
`Dataset` mirrors `Dataset::from_dense` (crates/lgbm-core/src/io/dataset.rs:904).
Errors surface as `BankaiError` carrying the library's last message, as the
reference's C API reports them through LGBM_GetLastError.

The library is required: importing this module does not load it, but every
call fails loudly when libbankai_gpu.so is missing. There is no CPU fallback.
"""
import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbankai_gpu.so")

FORM_PLAIN_F64 = 0
FORM_F32_COMPENSATED = 1
FORM_QUANTIZED = 2


class BankaiError(RuntimeError):
    pass


class DatasetView(C.Structure):
    _fields_ = [
        ("num_data", C.c_int64), ("num_groups", C.c_int32), ("num_features", C.c_int32),
        ("num_total_features", C.c_int32),
        ("group_bins", C.POINTER(C.POINTER(C.c_uint8))),
        ("group_num_total_bin", C.POINTER(C.c_int32)),
        ("group_feature_start", C.POINTER(C.c_int32)),
        ("group_feature_cnt", C.POINTER(C.c_int32)),
        ("group_bin_offsets", C.POINTER(C.c_int32)),
        ("feature_num_bin", C.POINTER(C.c_int32)),
        ("feature_missing_type", C.POINTER(C.c_int32)),
        ("feature_default_bin", C.POINTER(C.c_int32)),
        ("feature_most_freq_bin", C.POINTER(C.c_int32)),
        ("feature_bin_type", C.POINTER(C.c_int32)),
        ("feature_real_index", C.POINTER(C.c_int32)),
        ("feature_penalty", C.POINTER(C.c_double)),
        ("feature_bin_upper_bound", C.POINTER(C.POINTER(C.c_double))),
        ("feature_bin_2_categorical", C.POINTER(C.POINTER(C.c_int32))),
    ]


class LearnerParamsC(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int32), ("max_depth", C.c_int32), ("min_data_in_leaf", C.c_int32),
        ("min_sum_hessian_in_leaf", C.c_double), ("lambda_l1", C.c_double), ("lambda_l2", C.c_double),
        ("max_delta_step", C.c_double), ("path_smooth", C.c_double), ("min_gain_to_split", C.c_double),
        ("cat_smooth", C.c_double), ("cat_l2", C.c_double), ("max_cat_threshold", C.c_int32),
        ("max_cat_to_onehot", C.c_int32), ("min_data_per_group", C.c_int32), ("extra_trees", C.c_int32),
        ("histogram_pool_size", C.c_double), ("use_quantized_grad", C.c_int32),
        ("num_grad_quant_bins", C.c_int32), ("stochastic_rounding", C.c_int32),
        ("quant_train_renew_leaf", C.c_int32), ("seed", C.c_int32), ("num_threads", C.c_int32),
        ("form", C.c_int32), ("batched", C.c_int32), ("feature_fraction", C.c_double),
        ("feature_fraction_bynode", C.c_double), ("feature_fraction_seed", C.c_int32),
        ("num_interaction_constraints", C.c_int32), ("interaction_constraint_offsets", C.POINTER(C.c_int32)),
        ("interaction_constraint_features", C.POINTER(C.c_int32)),
    ]


class TreeOutC(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int32), ("num_cat", C.c_int32),
        ("split_feature_inner", C.POINTER(C.c_int32)), ("split_feature", C.POINTER(C.c_int32)),
        ("threshold_in_bin", C.POINTER(C.c_uint32)), ("threshold", C.POINTER(C.c_double)),
        ("decision_type", C.POINTER(C.c_int8)), ("left_child", C.POINTER(C.c_int32)),
        ("right_child", C.POINTER(C.c_int32)), ("split_gain", C.POINTER(C.c_float)),
        ("internal_value", C.POINTER(C.c_double)), ("internal_weight", C.POINTER(C.c_double)),
        ("internal_count", C.POINTER(C.c_int32)), ("leaf_value", C.POINTER(C.c_double)),
        ("leaf_weight", C.POINTER(C.c_double)), ("leaf_count", C.POINTER(C.c_int32)),
        ("leaf_parent", C.POINTER(C.c_int32)), ("leaf_depth", C.POINTER(C.c_int32)),
        ("cat_boundaries", C.POINTER(C.c_int32)), ("cat_threshold", C.POINTER(C.c_uint32)),
        ("cat_threshold_cap", C.c_int32), ("cat_threshold_len", C.c_int32),
    ]


class RankTablesC(C.Structure):
    _fields_ = [
        ("query_bounds", C.c_void_p), ("inverse_max_dcg", C.c_void_p), ("label_gain", C.c_void_p),
        ("discount", C.c_void_p), ("sigmoid_table", C.c_void_p),
        ("sigmoid", C.c_double), ("min_input", C.c_double), ("max_input", C.c_double), ("table_factor", C.c_double),
        ("num_queries", C.c_int32), ("num_gain", C.c_int32), ("table_bins", C.c_int32),
        ("truncation_level", C.c_int32), ("norm", C.c_int32), ("max_query_rows", C.c_int32),
    ]


def rank_tables_on_device(objective, device):
    """LambdaRank tables of a boosting.LambdarankObjective copied to `device` (torch):
    returns the RankTablesC and the tensors that back it (keep them alive)."""
    import torch

    keep = {
        "bounds": torch.as_tensor(np.asarray(objective.bounds, np.int32), device=device),
        "inv": torch.as_tensor(np.asarray(objective.inverse_max_dcg, np.float64), device=device),
        "gain": torch.as_tensor(np.asarray(objective.gain, np.float64), device=device),
        "discount": torch.as_tensor(np.asarray(objective.discount, np.float64), device=device),
        "table": torch.as_tensor(np.asarray(objective.table, np.float64), device=device),
    }
    t = RankTablesC()
    t.query_bounds = keep["bounds"].data_ptr()
    t.inverse_max_dcg = keep["inv"].data_ptr()
    t.label_gain = keep["gain"].data_ptr()
    t.discount = keep["discount"].data_ptr()
    t.sigmoid_table = keep["table"].data_ptr()
    t.sigmoid, t.min_input, t.max_input = objective.sigmoid, objective.min_in, objective.max_in
    t.table_factor = objective.factor
    t.num_queries = len(objective.bounds) - 1
    t.num_gain = len(objective.gain)
    t.table_bins = len(objective.table)
    t.truncation_level = objective.trunc
    t.norm = int(objective.norm)
    t.max_query_rows = int(np.max(np.diff(objective.bounds))) if len(objective.bounds) > 1 else 0
    return t, keep


class TrainStatsC(C.Structure):
    _fields_ = [("hist_rows", C.c_int64), ("part_rows", C.c_int64), ("splits", C.c_int32),
                ("rounds", C.c_int32), ("speculative_waste", C.c_int32), ("hist_bin_bytes", C.c_int32),
                ("committed_hist_rows", C.c_int64), ("committed_part_rows", C.c_int64),
                ("host_ms", C.c_double), ("wait_ms", C.c_double), ("hist_min_rows", C.c_int32),
                ("hist_max_blocks", C.c_int32), ("hist_tile_groups", C.c_int32), ("hist_ntiles", C.c_int32),
                ("hist_sub_histograms", C.c_int32), ("partition_histograms", C.c_int32),
                ("replay_ms", C.c_double), ("plan_ms", C.c_double), ("launch_ms", C.c_double),
                ("readback_ms", C.c_double)]


class KernelStatsC(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_int64), ("algo_bytes", C.c_double), ("moved_bytes", C.c_double)]


KERNEL_KINDS = ("root_hist", "child_hist", "partition", "search", "discretize", "renew", "score_gradients", "plan",
                "transfers")


class DatasetParamsC(C.Structure):
    _fields_ = [("max_bin", C.c_int32), ("min_data_in_bin", C.c_int32),
                ("bin_construct_sample_cnt", C.c_int32), ("data_random_seed", C.c_int32),
                ("use_missing", C.c_int32), ("zero_as_missing", C.c_int32),
                ("feature_pre_filter", C.c_int32), ("min_data_in_leaf", C.c_int32),
                ("num_threads", C.c_int32), ("enable_bundle", C.c_int32), ("is_enable_sparse", C.c_int32),
                ("gpu_bundle_limit", C.c_int32), ("unbundle_groups", C.c_int32)]


# Every entry point the headers declare, with its argument types.
_P = C.c_void_p
EXPORTS = {
    "bk_last_error": ([], C.c_char_p),
    "bk_learner_create": ([C.POINTER(DatasetView), C.POINTER(LearnerParamsC), C.c_int, C.POINTER(_P)], C.c_int),
    "bk_learner_train": ([_P, _P, _P, _P, C.c_int64, C.POINTER(TreeOutC)], C.c_int),
    "bk_learner_train_device": ([_P, _P, _P, C.POINTER(TreeOutC)], C.c_int),
    "bk_learner_train_device_bagged": ([_P, _P, _P, _P, C.c_int64, C.POINTER(TreeOutC)], C.c_int),
    "bk_learner_add_out_of_bag_prediction": ([_P, C.POINTER(TreeOutC), _P, C.c_int64, _P], C.c_int),
    "bk_learner_add_out_of_bag_prediction_device": ([_P, C.POINTER(TreeOutC), _P], C.c_int),
    "bk_learner_last_splits": ([_P, _P, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "bk_learner_histogram_cache_capacity": ([_P, C.POINTER(C.c_int64)], C.c_int),
    "bk_learner_reset_histogram_pool_config": ([_P], C.c_int),
    "bk_learner_continue_after_dataset_reset": ([_P, _P], C.c_int),
    "bk_learner_add_prediction_for_rows": ([_P, C.POINTER(TreeOutC), _P, C.c_int32, _P, C.c_int64, _P], C.c_int),
    "bk_learner_add_prediction_to_score": ([_P, C.POINTER(TreeOutC), _P], C.c_int),
    "bk_learner_add_prediction_to_score_device": ([_P, C.POINTER(TreeOutC), _P], C.c_int),
    "bk_learner_leaf_rows": ([_P, C.c_int32, _P, C.POINTER(C.c_int64)], C.c_int),
    "bk_learner_reset_constant_hessian": ([_P, C.c_int], C.c_int),
    "bk_learner_binary_gradients_device": ([_P, _P, _P, _P, C.c_double, _P, _P, _P], C.c_int),
    "bk_learner_softmax_gradients_device": ([_P, _P, _P, C.c_int, _P, _P], C.c_int),
    "bk_learner_set_tree_slot": ([_P, C.c_int, C.c_int], C.c_int),
    "bk_learner_lambdarank_gradients_device": ([_P, _P, _P, _P, _P, _P], C.c_int),
    "bk_learner_last_stats": ([_P, C.POINTER(TrainStatsC)], C.c_int),
    "bk_learner_profile": ([_P, C.c_int], C.c_int),
    "bk_learner_kernel_stats": ([_P, C.c_int, C.POINTER(KernelStatsC)], C.c_int),
    "bk_learner_root_histogram": ([_P, _P, _P, _P, _P, _P], C.c_int),
    "bk_learner_profile_trees": ([_P, C.c_int], C.c_int),
    "bk_learner_set_hist_tiling": ([_P, C.c_int32, C.c_int32], C.c_int),
    "bk_learner_set_hist_bin_width": ([_P, C.c_int32], C.c_int),
    "bk_learner_set_hist_layout": ([_P, C.c_int32, C.c_int32], C.c_int),
    "bk_learner_set_partition_histograms": ([_P, C.c_int32], C.c_int),
    "bk_learner_set_round_planner": ([_P, C.c_int], C.c_int),
    "bk_learner_root_histogram_noalloc": ([_P, _P, _P, _P, _P], C.c_int),
    "bk_learner_free": ([_P], C.c_int),
    "bk_dataset_from_dense": ([_P, C.c_int64, C.c_int32, C.POINTER(DatasetParamsC), _P, C.c_int32, C.POINTER(_P)],
                              C.c_int),
    "bk_dataset_from_dense_f32": ([_P, C.c_int64, C.c_int32, C.POINTER(DatasetParamsC), _P, C.c_int32,
                                   C.POINTER(_P)], C.c_int),
    "bk_dataset_from_dense_f32_reference": ([_P, C.c_int64, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "bk_dataset_get_view": ([_P, C.POINTER(DatasetView)], C.c_int),
    "bk_dataset_feature_info": ([_P, C.c_int32, C.c_char_p, C.c_int32], C.c_int),
    "bk_dataset_free": ([_P], C.c_int),
    "bk_dataset_bundle_layout": ([_P, C.c_int64, C.c_int32, C.POINTER(DatasetParamsC), _P, C.c_int32, _P, _P, _P, _P,
                                  C.POINTER(C.c_int32)], C.c_int),
}

_lib = None


def load_library(path=LIB_PATH):
    """Loads libbankai_gpu.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BankaiError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (args, res) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise BankaiError(_lib.bk_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _score_ptr(score, num_data):
    """A score array the library writes float64 values into: it must be a
    C-contiguous, writeable float64 numpy array of at least num_data entries."""
    if not isinstance(score, np.ndarray) or score.dtype != np.float64 or not score.flags["C_CONTIGUOUS"] \
            or not score.flags["WRITEABLE"] or score.ndim != 1 or score.shape[0] < num_data:
        raise BankaiError(f"score must be a writeable C-contiguous float64 array of at least {num_data} entries")
    return _ptr(score)


def _rows_array(rows, num_data):
    """Row ids as C-contiguous int32, each inside [0, num_data)."""
    r = np.ascontiguousarray(rows, dtype=np.int32)
    if r.ndim != 1 or (len(r) and (int(r.min()) < 0 or int(r.max()) >= num_data)):
        raise BankaiError("rows must be a 1-d array of row ids inside the dataset")
    return r


@dataclass
class DatasetParams:
    """Dataset-construction config; defaults are the reference's (config/auto.rs:244-251)."""
    max_bin: int = 255
    min_data_in_bin: int = 3
    bin_construct_sample_cnt: int = 200000
    data_random_seed: int = 1
    use_missing: int = 1
    zero_as_missing: int = 0
    feature_pre_filter: int = 1
    min_data_in_leaf: int = 20
    num_threads: int = 0
    enable_bundle: int = 1
    is_enable_sparse: int = 1
    gpu_bundle_limit: int = 0
    unbundle_groups: int = 0

    def to_c(self):
        return DatasetParamsC(**{f.name: getattr(self, f.name) for f in fields(self)})


def bundle_layout(data, params=None, categorical=()):
    """The group layout Dataset::Construct gives `data` (bk_dataset_bundle_layout):
    per column its group and sub-feature index (-1 when unused), per group its bins
    and whether it is multi-valued; also for layouts the dataset cannot store."""
    lib = load_library()
    x = np.ascontiguousarray(data, dtype=np.float64)
    nrow, ncol = x.shape
    p = (params or DatasetParams()).to_c()
    cats = np.asarray(list(categorical), dtype=np.int32)
    fg = np.zeros(ncol, np.int32)
    fs = np.zeros(ncol, np.int32)
    gb = np.zeros(ncol, np.int64)
    gm = np.zeros(ncol, np.int8)
    ng = C.c_int32()
    _check(lib.bk_dataset_bundle_layout(_ptr(x), nrow, ncol, C.byref(p), _ptr(cats) if len(cats) else None, len(cats),
                                        _ptr(fg), _ptr(fs), _ptr(gb), _ptr(gm), C.byref(ng)))
    n = ng.value
    return {"feature_group": fg, "feature_sub": fs, "group_num_total_bin": gb[:n].copy(),
            "group_multi_val": gm[:n].astype(bool)}


def derive_data_random_seed(seed):
    """Config::set derives data_random_seed from `seed` (config/mod.rs:397-405)."""
    x = seed & 0xffffffff
    x = (x * 214013 + 2531011) & 0xffffffff
    return ((x >> 16) & 0x7fff) % 32767


class Dataset:
    """A binned dataset (Dataset::from_dense). `data` is a row-major (nrow, ncol) array."""

    def __init__(self, data=None, params=None, categorical=(), _handle=None, _keep=None):
        lib = load_library()
        self._lib = lib
        if _handle is not None:
            self.handle = _handle
            self._keep = _keep
        else:
            params = params or DatasetParams()
            data = np.ascontiguousarray(data)
            nrow, ncol = data.shape
            cats = np.asarray(list(categorical), dtype=np.int32)
            out = C.c_void_p()
            p = params.to_c()
            if data.dtype == np.float32:
                _check(lib.bk_dataset_from_dense_f32(_ptr(data), nrow, ncol, C.byref(p), _ptr(cats), len(cats),
                                                     C.byref(out)))
            else:
                data = np.ascontiguousarray(data, dtype=np.float64)
                _check(lib.bk_dataset_from_dense(_ptr(data), nrow, ncol, C.byref(p), _ptr(cats), len(cats),
                                                 C.byref(out)))
            self.handle = out
        self.view = DatasetView()
        _check(lib.bk_dataset_get_view(self.handle, C.byref(self.view)))

    def bin_like(self, data):
        """Bins new rows with this dataset's mappers (Dataset::from_dense_reference)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        out = C.c_void_p()
        _check(self._lib.bk_dataset_from_dense_f32_reference(_ptr(data), data.shape[0], data.shape[1], self.handle,
                                                             C.byref(out)))
        return Dataset(_handle=out)

    @property
    def num_data(self):
        return self.view.num_data

    @property
    def num_features(self):
        return self.view.num_features

    def feature_infos(self):
        buf = C.create_string_buffer(1 << 16)
        out = []
        for c in range(self.view.num_total_features):
            _check(self._lib.bk_dataset_feature_info(self.handle, c, buf, len(buf)))
            out.append(buf.value.decode())
        return out

    def group_columns(self):
        """The stored u8 column of every group, as numpy arrays (copies)."""
        v = self.view
        return [np.ctypeslib.as_array(v.group_bins[g], shape=(v.num_data,)).copy() for g in range(v.num_groups)]

    def group_num_total_bin(self):
        """Bins of every group; the learner's histograms concatenate the groups."""
        v = self.view
        return np.ctypeslib.as_array(v.group_num_total_bin, shape=(v.num_groups,)).copy()

    def feature_meta(self):
        v = self.view
        nf = v.num_features
        arr = lambda p: np.ctypeslib.as_array(p, shape=(nf,)).copy()  # noqa: E731
        return {
            "num_bin": arr(v.feature_num_bin), "missing_type": arr(v.feature_missing_type),
            "default_bin": arr(v.feature_default_bin), "most_freq_bin": arr(v.feature_most_freq_bin),
            "bin_type": arr(v.feature_bin_type), "real_index": arr(v.feature_real_index),
        }

    def __del__(self):
        if getattr(self, "handle", None) and self._lib is not None:
            self._lib.bk_dataset_free(self.handle)
            self.handle = None


@dataclass
class LearnerParams:
    """Tree-learner config; defaults are the reference's (config/auto.rs:163-251)."""
    num_leaves: int = 31
    max_depth: int = -1
    min_data_in_leaf: int = 20
    min_sum_hessian_in_leaf: float = 1e-3
    lambda_l1: float = 0.0
    lambda_l2: float = 0.0
    max_delta_step: float = 0.0
    path_smooth: float = 0.0
    min_gain_to_split: float = 0.0
    cat_smooth: float = 10.0
    cat_l2: float = 10.0
    max_cat_threshold: int = 32
    max_cat_to_onehot: int = 4
    min_data_per_group: int = 100
    extra_trees: int = 0
    histogram_pool_size: float = -1.0
    use_quantized_grad: int = 0
    num_grad_quant_bins: int = 4
    stochastic_rounding: int = 1
    quant_train_renew_leaf: int = 0
    seed: int = 0
    num_threads: int = 1
    form: int = FORM_PLAIN_F64
    batched: int = 0
    feature_fraction: float = 1.0
    feature_fraction_bynode: float = 1.0
    feature_fraction_seed: int = 2
    interaction_constraints: tuple = ()  # lists of real feature ids

    def to_c(self):
        kw = {f.name: getattr(self, f.name) for f in fields(self) if f.name != "interaction_constraints"}
        c = LearnerParamsC(**kw)
        groups = [list(g) for g in self.interaction_constraints]
        offsets = np.zeros(len(groups) + 1, np.int32)
        for i, g in enumerate(groups):
            offsets[i + 1] = offsets[i] + len(g)
        feats = np.asarray([f for g in groups for f in g] or [0], np.int32)
        c.num_interaction_constraints = len(groups)
        c.interaction_constraint_offsets = offsets.ctypes.data_as(C.POINTER(C.c_int32))
        c.interaction_constraint_features = feats.ctypes.data_as(C.POINTER(C.c_int32))
        c._keep = (offsets, feats)  # the arrays live as long as the struct
        return c


SPLIT_RECORD_BYTES = 128  # sizeof(bk_split_record)


class TreeBuffers:
    """Host arrays a bk_tree_out points into; `as_dict()` returns trimmed copies."""

    @classmethod
    def from_dict(cls, d, num_leaves_cap):
        """A bk_tree_out over copies of a tree dict (as train() returns it)."""
        tb = cls(num_leaves_cap, cat_cap=max(len(d["cat_threshold"]), 1))
        for k, v in tb.a.items():
            src = np.asarray(d[k])
            v[: len(src)] = src
        tb.c.num_leaves = d["num_leaves"]
        tb.c.num_cat = d["num_cat"]
        tb.c.cat_threshold_len = len(d["cat_threshold"])
        return tb

    def __init__(self, num_leaves, cat_cap=1 << 16):
        ni, nl = max(num_leaves - 1, 1), num_leaves
        self.a = {
            "split_feature_inner": np.zeros(ni, np.int32), "split_feature": np.zeros(ni, np.int32),
            "threshold_in_bin": np.zeros(ni, np.uint32), "threshold": np.zeros(ni, np.float64),
            "decision_type": np.zeros(ni, np.int8), "left_child": np.zeros(ni, np.int32),
            "right_child": np.zeros(ni, np.int32), "split_gain": np.zeros(ni, np.float32),
            "internal_value": np.zeros(ni, np.float64), "internal_weight": np.zeros(ni, np.float64),
            "internal_count": np.zeros(ni, np.int32), "leaf_value": np.zeros(nl, np.float64),
            "leaf_weight": np.zeros(nl, np.float64), "leaf_count": np.zeros(nl, np.int32),
            "leaf_parent": np.zeros(nl, np.int32), "leaf_depth": np.zeros(nl, np.int32),
            "cat_boundaries": np.zeros(nl, np.int32), "cat_threshold": np.zeros(cat_cap, np.uint32),
        }
        ctypes_of = {np.int32: C.c_int32, np.uint32: C.c_uint32, np.float64: C.c_double, np.int8: C.c_int8,
                     np.float32: C.c_float}
        kw = {k: v.ctypes.data_as(C.POINTER(ctypes_of[v.dtype.type])) for k, v in self.a.items()}
        self.c = TreeOutC(num_leaves=1, num_cat=0, cat_threshold_cap=cat_cap, cat_threshold_len=0, **kw)

    def as_dict(self):
        nl = self.c.num_leaves
        ni = nl - 1
        d = {"num_leaves": nl, "num_cat": self.c.num_cat}
        for k, v in self.a.items():
            if k in ("leaf_value", "leaf_weight", "leaf_count", "leaf_parent", "leaf_depth"):
                d[k] = v[:nl].copy()
            elif k == "cat_boundaries":
                d[k] = v[:self.c.num_cat + 1].copy()
            elif k == "cat_threshold":
                d[k] = v[:self.c.cat_threshold_len].copy()
            else:
                d[k] = v[:ni].copy()
        return d

    def set_leaf_values(self, values):
        self.a["leaf_value"][: len(values)] = values


class Learner:
    """The GPU tree learner behind the SerialTreeLearner seat."""

    def __init__(self, dataset, params=None, is_constant_hessian=False):
        self._lib = load_library()
        self.params = params or LearnerParams()
        self.dataset = dataset
        p = self.params.to_c()
        out = C.c_void_p()
        _check(self._lib.bk_learner_create(C.byref(dataset.view), C.byref(p), int(is_constant_hessian),
                                           C.byref(out)))
        self.handle = out
        self.tree = TreeBuffers(self.params.num_leaves)

    def train(self, gradients, hessians, rows=None):
        """SerialTreeLearner::train on host arrays; returns the tree as a dict.
        rows: the bag (ascending row ids) the tree is grown on, or None for all rows."""
        g = np.ascontiguousarray(gradients, dtype=np.float32)
        h = np.ascontiguousarray(hessians, dtype=np.float32)
        if g.shape[0] != self.dataset.num_data or h.shape[0] != self.dataset.num_data:
            raise BankaiError("gradients and hessians must have num_data entries")
        r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
        _check(self._lib.bk_learner_train(self.handle, _ptr(g), _ptr(h), None if r is None else _ptr(r),
                                          0 if r is None else len(r), C.byref(self.tree.c)))
        return self.tree.as_dict()

    def train_device_bagged(self, d_gradients_ptr, d_hessians_ptr, d_rows_ptr, num_rows):
        """train() on device arrays with a device bag of num_rows ascending int32 rows."""
        _check(self._lib.bk_learner_train_device_bagged(self.handle, C.c_void_p(d_gradients_ptr),
                                                        C.c_void_p(d_hessians_ptr), C.c_void_p(d_rows_ptr),
                                                        int(num_rows), C.byref(self.tree.c)))
        return self.tree.as_dict()

    def add_out_of_bag_prediction(self, leaf_values, rows, score):
        """score[row] += leaf_values[leaf(row)] for the given rows (host float64 array)."""
        self.tree.set_leaf_values(leaf_values)
        r = _rows_array(rows, self.dataset.num_data)
        sp = _score_ptr(score, self.dataset.num_data)
        _check(self._lib.bk_learner_add_out_of_bag_prediction(self.handle, C.byref(self.tree.c), _ptr(r), len(r),
                                                              sp))

    def histogram_cache_capacity(self):
        out = C.c_int64()
        _check(self._lib.bk_learner_histogram_cache_capacity(self.handle, C.byref(out)))
        return out.value

    def reset_histogram_pool_config(self):
        _check(self._lib.bk_learner_reset_histogram_pool_config(self.handle))

    def continue_after_dataset_reset(self, previous):
        _check(self._lib.bk_learner_continue_after_dataset_reset(self.handle, previous.handle))

    def last_splits(self):
        """The last tree's split records (an opaque uint8 array, one row per internal node)."""
        n = C.c_int32()
        _check(self._lib.bk_learner_last_splits(self.handle, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), SPLIT_RECORD_BYTES), np.uint8)
        _check(self._lib.bk_learner_last_splits(self.handle, _ptr(out), n.value, C.byref(n)))
        return out[: n.value].copy()

    def add_prediction_for_rows(self, tree, leaf_values, splits, rows, score):
        """An earlier tree (the dict train() returned, with leaf_values in place of its
        own) replayed on rows through its split records; score[row] += its leaf value."""
        d = dict(tree)
        d["leaf_value"] = np.asarray(leaf_values, np.float64)
        tb = TreeBuffers.from_dict(d, self.params.num_leaves)
        sp = np.ascontiguousarray(splits, np.uint8)
        r = _rows_array(rows, self.dataset.num_data)
        dst = _score_ptr(score, self.dataset.num_data)
        _check(self._lib.bk_learner_add_prediction_for_rows(self.handle, C.byref(tb.c), _ptr(sp), len(sp), _ptr(r),
                                                            len(r), dst))

    def add_out_of_bag_prediction_device(self, leaf_values, d_score_ptr):
        self.tree.set_leaf_values(leaf_values)
        _check(self._lib.bk_learner_add_out_of_bag_prediction_device(self.handle, C.byref(self.tree.c),
                                                                     C.c_void_p(d_score_ptr)))

    def train_device(self, d_gradients_ptr, d_hessians_ptr):
        """train() on device arrays given as raw pointers (e.g. tensor.data_ptr())."""
        _check(self._lib.bk_learner_train_device(self.handle, C.c_void_p(d_gradients_ptr),
                                                 C.c_void_p(d_hessians_ptr), C.byref(self.tree.c)))
        return self.tree.as_dict()

    def add_prediction_to_score(self, leaf_values, score):
        """score[row] += leaf_values[leaf(row)] for the last tree (host float64 array, in place)."""
        self.tree.set_leaf_values(leaf_values)
        sp = _score_ptr(score, self.dataset.num_data)
        _check(self._lib.bk_learner_add_prediction_to_score(self.handle, C.byref(self.tree.c), sp))

    def add_prediction_to_score_device(self, leaf_values, d_score_ptr):
        self.tree.set_leaf_values(leaf_values)
        _check(self._lib.bk_learner_add_prediction_to_score_device(self.handle, C.byref(self.tree.c),
                                                                   C.c_void_p(d_score_ptr)))

    def binary_gradients_device(self, d_score_ptr, d_label_ptr, sigmoid, d_g_ptr, d_h_ptr, label_weights=(1.0, 1.0),
                                d_weights_ptr=None):
        """BinaryLogloss gradients on the device: label_weights = (negative, positive)
        as the objective derives them (BinaryObjective.label_weights); optional
        per-row sample weights (a device float32 pointer)."""
        lw = (C.c_double * 2)(float(label_weights[0]), float(label_weights[1]))
        _check(self._lib.bk_learner_binary_gradients_device(self.handle, C.c_void_p(d_score_ptr),
                                                            C.c_void_p(d_label_ptr),
                                                            C.c_void_p(d_weights_ptr) if d_weights_ptr else None,
                                                            float(sigmoid), lw, C.c_void_p(d_g_ptr),
                                                            C.c_void_p(d_h_ptr)))

    def lambdarank_gradients_device(self, d_score_ptr, d_label_ptr, tables, d_g_ptr, d_h_ptr):
        """tables: a RankTablesC whose arrays are device pointers."""
        _check(self._lib.bk_learner_lambdarank_gradients_device(self.handle, C.c_void_p(d_score_ptr),
                                                                C.c_void_p(d_label_ptr), C.byref(tables),
                                                                C.c_void_p(d_g_ptr), C.c_void_p(d_h_ptr)))

    def set_tree_slot(self, slot, trees_per_iteration):
        """Train tree `slot` of every `trees_per_iteration` trees of a shared sequence."""
        _check(self._lib.bk_learner_set_tree_slot(self.handle, int(slot), int(trees_per_iteration)))

    def softmax_gradients_device(self, d_score_ptr, d_label_ptr, num_class, d_g_ptr, d_h_ptr):
        _check(self._lib.bk_learner_softmax_gradients_device(self.handle, C.c_void_p(d_score_ptr),
                                                             C.c_void_p(d_label_ptr), int(num_class),
                                                             C.c_void_p(d_g_ptr), C.c_void_p(d_h_ptr)))

    def leaf_rows(self, leaf):
        out = np.zeros(self.dataset.num_data, np.int32)
        n = C.c_int64()
        _check(self._lib.bk_learner_leaf_rows(self.handle, leaf, _ptr(out), C.byref(n)))
        return out[: n.value].copy()

    def root_histogram(self, gradients, hessians):
        """Root histogram before fix_histogram: (gh, counts), where gh is a
        (total_bins, 2) float64 array for the float forms and int64 for the
        quantized form (exact integer sums of the discretized gradients)."""
        g = np.ascontiguousarray(gradients, dtype=np.float32)
        h = np.ascontiguousarray(hessians, dtype=np.float32)
        if g.shape[0] != self.dataset.num_data or h.shape[0] != self.dataset.num_data:
            raise BankaiError("gradients and hessians must have num_data entries")
        tb = int(np.sum(self.dataset.group_num_total_bin()))
        cnt = np.zeros(tb, np.int32)
        if self.params.form == FORM_QUANTIZED:
            gh = np.zeros((tb, 2), np.int64)
            _check(self._lib.bk_learner_root_histogram(self.handle, _ptr(g), _ptr(h), None, _ptr(gh), _ptr(cnt)))
        else:
            gh = np.zeros((tb, 2), np.float64)
            _check(self._lib.bk_learner_root_histogram(self.handle, _ptr(g), _ptr(h), _ptr(gh), None, _ptr(cnt)))
        return gh, cnt

    def root_histogram_noalloc(self, d_gradients_ptr, d_hessians_ptr, d_hist_ptr, d_count_ptr):
        """bk_learner_root_histogram_noalloc: device inputs and outputs (raw pointers)."""
        _check(self._lib.bk_learner_root_histogram_noalloc(self.handle, C.c_void_p(d_gradients_ptr),
                                                           C.c_void_p(d_hessians_ptr), C.c_void_p(d_hist_ptr),
                                                           C.c_void_p(d_count_ptr)))

    def set_round_planner(self, mode):
        """Batched growth planned on the host (0, the default), on the device for
        the whole tree (1), or on the device for the rounds only (2)."""
        _check(self._lib.bk_learner_set_round_planner(self.handle, int(mode)))

    def set_hist_tiling(self, min_rows=0, max_blocks=0):
        """Forces the histogram tiling (0, 0: the autotuner's choice)."""
        _check(self._lib.bk_learner_set_hist_tiling(self.handle, int(min_rows), int(max_blocks)))

    def set_hist_layout(self, groups_per_tile=0, sub_histograms=0):
        """Forces the histogram layout: a new LDS tile every groups_per_tile groups
        (a multiple of 16) and 1 or 2 LDS copies per tile (0, 0: the autotuner's)."""
        _check(self._lib.bk_learner_set_hist_layout(self.handle, int(groups_per_tile), int(sub_histograms)))

    def set_partition_histograms(self, on=True):
        """With 4-byte bins: the partition also builds one child's histogram per
        split (True) or a separate launch builds the smaller child's (False, the
        default, measured faster)."""
        _check(self._lib.bk_learner_set_partition_histograms(self.handle, 1 if on else 0))

    def set_hist_bin_width(self, nbytes=0):
        """Bytes per quantized histogram bin of batched growth: 4 (no row counts;
        the partition counts the children), 8 (with counts), 0 (4 where it applies)."""
        _check(self._lib.bk_learner_set_hist_bin_width(self.handle, int(nbytes)))

    def reset_constant_hessian(self, value):
        _check(self._lib.bk_learner_reset_constant_hessian(self.handle, int(value)))

    def last_stats(self):
        s = TrainStatsC()
        _check(self._lib.bk_learner_last_stats(self.handle, C.byref(s)))
        return {"hist_rows": s.hist_rows, "part_rows": s.part_rows, "splits": s.splits, "rounds": s.rounds,
                "committed_hist_rows": s.committed_hist_rows, "committed_part_rows": s.committed_part_rows,
                "host_ms": s.host_ms, "wait_ms": s.wait_ms,
                "speculative_waste": s.speculative_waste, "hist_bin_bytes": s.hist_bin_bytes,
                "partition_histograms": s.partition_histograms,
                "host_phases_ms": {"replay": s.replay_ms, "plan": s.plan_ms, "launch": s.launch_ms,
                                   "readback": s.readback_ms},
                "hist_tiling": {"min_rows": s.hist_min_rows, "max_blocks": s.hist_max_blocks,
                                "tile_groups": s.hist_tile_groups, "ntiles": s.hist_ntiles,
                                "sub_histograms": s.hist_sub_histograms}}

    def profile(self, kinds=True, every_trees=1):
        """Per-kernel HIP-event timing: True for every kind, False for none, or a
        list of names from KERNEL_KINDS, on the launches of every `every_trees`-th
        tree. Resets the counters."""
        if kinds is True:
            mask = (1 << len(KERNEL_KINDS)) - 1
        elif not kinds:
            mask = 0
        else:
            mask = sum(1 << KERNEL_KINDS.index(k) for k in kinds)
        _check(self._lib.bk_learner_profile(self.handle, mask))
        _check(self._lib.bk_learner_profile_trees(self.handle, int(every_trees)))

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(KERNEL_KINDS):
            s = KernelStatsC()
            _check(self._lib.bk_learner_kernel_stats(self.handle, k, C.byref(s)))
            out[name] = {"ms": s.ms, "launches": s.launches, "algo_bytes": s.algo_bytes, "moved_bytes": s.moved_bytes}
        return out

    def close(self):
        if getattr(self, "handle", None):
            self._lib.bk_learner_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
