"""Test configuration: the `gpu` marker, import paths, and the shared loaders.

Tests marked `gpu` run the HIP library on a device; everything else runs on the
CPU (the oracle, host logic, and that the C-ABI library loads and exports its
symbols).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bankai-lightgbm_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libbankai_gpu.so")


def pytest_collection_modifyitems(config, items):
    # GPU tests are collected everywhere but skipped when no device is visible
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover - torch is part of the image
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
