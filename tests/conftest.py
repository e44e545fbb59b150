"""Test configuration: import paths and the ``gpu`` marker.

``-m "not gpu"`` runs everything that needs no GPU (the oracle against the
reference's published numbers and closed forms, host logic, the C-ABI exports);
``-m gpu`` runs the parity tests of the HIP kernels against the oracle.
"""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "visual-computing_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libgdpoisson.so")
