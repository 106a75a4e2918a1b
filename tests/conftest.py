"""Shared fixtures: the product package (loaded from its directory), the oracle (checker), the
golden fixtures of the reference's v0, and the `gpu` marker."""
from __future__ import annotations

import json
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def lbm():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


def load_golden(name: str):
    z = np.load(GOLDEN / name, allow_pickle=False)
    return z["solid"], z["S"], json.loads(str(z["params"]))


def walls_block_mask(nx=53, ny=31):
    m = np.zeros((ny, nx), np.uint8)
    m[0, :] = 1
    m[-1, :] = 1
    m[10:17, 12:16] = 1
    return m.reshape(-1)
