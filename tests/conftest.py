import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")

# the sanitizer run's copy of the tree (tests/sanitizer/run.sh) is not collected here
collect_ignore_glob = ["sanitizer/_build/*"]
