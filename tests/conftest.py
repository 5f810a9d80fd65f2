import os
import sys
from pathlib import Path

import pytest

# torch ships its own HIP runtime next to the one the library links (/opt/rocm). Loaded after
# the library has initialised its runtime, torch's finds no GPU; loaded first, both work (the
# order bench.py uses), so it is imported before any test can touch the library.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "hnsw-engine_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

try:
    from hypothesis import HealthCheck, settings

    settings.register_profile(
        "default",
        deadline=None,
        max_examples=40,
        suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture],
    )
    settings.load_profile(os.environ.get("HYPOTHESIS_PROFILE", "default"))
except ImportError:  # pragma: no cover
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
