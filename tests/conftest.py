import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built liblmm_hip.so")


def pytest_runtest_logstart(nodeid, location):
    # one flushed line per test so a log tail names the test that was running
    print(f"\nSTART {nodeid}", flush=True)
