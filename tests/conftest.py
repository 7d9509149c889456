"""Test configuration: import paths and the ``gpu`` marker.

``-m "not gpu"`` runs on any host (oracle, host logic, library exports);
``-m gpu`` needs an MI355X and calls the HIP engine through the C ABI.
"""
from __future__ import annotations

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "gpuwm-global_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")
