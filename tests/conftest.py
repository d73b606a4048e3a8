import os
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The launch-shape autotuner's cache goes to a fresh directory per test run
# (not the user's ~/.cache), so every run tunes from scratch.
os.environ.setdefault("PYSA_AMD_TUNE_CACHE",
                      os.path.join(tempfile.mkdtemp(prefix="pysa_amd_tune_"), "tune.json"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")
