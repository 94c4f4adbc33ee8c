"""Shared test set-up: import paths, the `gpu` marker, golden vectors."""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "engel-cuda-negacyclic-multiply_amd"
for p in (ROOT / "oracle", PKG, ROOT / "tests" / "golden"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN_DIR = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_golden():
    return json.loads((GOLDEN_DIR / "golden.json").read_text())


def modulus(name: str) -> np.ndarray:
    doc = load_golden()["moduli"][name]
    L = doc["L"]
    return np.frombuffer(int(doc["hex"], 16).to_bytes(4 * L, "little"), dtype="<u4").astype(np.uint32)


@pytest.fixture(scope="session")
def golden():
    return load_golden()


@pytest.fixture(scope="session")
def q868():
    return modulus("q868")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no HIP device")
    import torch

    torch.cuda.init()
    return torch.device("cuda:0")
