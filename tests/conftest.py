"""Shared test set-up.

The product package lives in a directory whose name is not a Python
identifier, so it is imported with ``importlib`` and registered under the
alias ``pfflip_amd``.  GPU tests are marked ``gpu``; everything else runs on
the CPU.
"""

import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_DIR = "implementation-of-adaptive-phase-field-flip_amd"


def load_pkg():
    mod = importlib.import_module(PKG_DIR)
    sys.modules.setdefault("pfflip_amd", mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def pf():
    return load_pkg()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


_FAULT_MARKERS = ("illegal memory access", "AcceleratorError", "hipErrorIllegalAddress",
                  "memory access fault", "HSA_STATUS_ERROR")


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    """Stop the whole session at the first GPU fault: every later GPU call
    would fail anyway, and nothing more should run on a faulted device."""
    outcome = yield
    rep = outcome.get_result()
    if rep.failed and call.excinfo is not None:
        text = f"{call.excinfo.typename}: {call.excinfo.value}"
        if any(m in text for m in _FAULT_MARKERS):
            item.session.shouldstop = f"GPU fault in {item.nodeid}; stopping"
