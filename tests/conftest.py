"""pytest setup: import paths for the package and the oracle, markers."""
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent
for sub in ("coldplate-solver_amd", "oracle"):
    p = str(REPO / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
