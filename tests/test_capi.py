"""CPU checks of the C-ABI library: it loads, exports every function its headers
declare, builds datasets on the host, and reports errors through bk_last_error()
(crates/lgbm-capi/src/lib.rs:63-90,181 convention) without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bankai_gpu as bg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("bankai_gpu.h", "bankai_dataset.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(bk_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = bg.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(bg.EXPORTS)


def test_dataset_construction_on_host():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5000, 6)).astype(np.float32)
    x[:, 3] = rng.integers(0, 5, size=5000)
    ds = bg.Dataset(x, categorical=[3])
    assert ds.num_data == 5000
    meta = ds.feature_meta()
    assert sorted(meta["real_index"].tolist()) == list(range(6))
    infos = ds.feature_infos()
    assert infos[3].startswith("-1:")
    cols = ds.group_columns()
    assert all(c.dtype == np.uint8 and len(c) == 5000 for c in cols)


def test_errors_are_reported_not_raised_across_the_abi():
    lib = bg.load_library()
    out = C.c_void_p()
    p = bg.DatasetParams().to_c()
    rc = lib.bk_dataset_from_dense(None, 0, 3, C.byref(p), None, 0, C.byref(out))
    assert rc == -1
    assert b"empty" in lib.bk_last_error()


def test_unsupported_parameters_are_refused():
    import torch
    if torch.cuda.is_available():
        pytest.skip("checks the device-less error path")
    x = np.random.default_rng(1).normal(size=(100, 3)).astype(np.float32)
    ds = bg.Dataset(x)
    with pytest.raises(bg.BankaiError):
        bg.Learner(ds, bg.LearnerParams(num_leaves=4))
