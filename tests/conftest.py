import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def pytest_sessionfinish(session, exitstatus):
    # release every device handle the tests left open before interpreter
    # shutdown (solvers, batches, lattices, dense factors: streams, device
    # memory and rocBLAS handles)
    mod = sys.modules.get("aeroflex_amd._handle")
    if mod is not None:
        mod.close_all()
