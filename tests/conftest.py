"""Shared fixtures.  `gpu` marks tests that need an MI355X; everything else runs on CPU."""
from __future__ import annotations

import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu on the GPU host)")


def _ensure_built():
    need = [
        os.path.join(ROOT, "pars_amd", "libpars_hip.so"),
        os.path.join(ROOT, "pars_amd", "libpars_gen.so"),
        os.path.join(ROOT, "oracle", "liboracle.so"),
    ]
    if not all(os.path.exists(p) for p in need):
        from pars_amd import build

        build.build_all()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def torch_cuda():
    """torch on the GPU (imported before the backend so one HIP runtime serves both)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    return torch
