import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libifs.so")


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_protocol(item, nextitem):
    """Names each GPU test on the terminal as it starts when the run is quiet (-q), whose
    only other output is one character per finished test: a long GPU test is then
    visible while it runs, not only once it ends."""
    tr = item.config.pluginmanager.getplugin("terminalreporter")
    if tr is not None and item.config.getoption("verbose") <= 0 and item.get_closest_marker("gpu"):
        tr.write_line(f"start {item.nodeid}")
        tr.flush()
    yield


@pytest.fixture
def progress(request):
    """progress(msg): one line on the terminal at once, past output capture, with the
    seconds since the test began.  Long tests report each finished step with it, so a
    run without -s shows their steps as they finish."""
    import time
    pm = request.config.pluginmanager
    capman, tr = pm.getplugin("capturemanager"), pm.getplugin("terminalreporter")
    t0 = time.perf_counter()

    def say(msg):
        line = f"[{request.node.name} {time.perf_counter() - t0:.1f} s] {msg}"
        if capman is None or tr is None:
            print(line, flush=True)
            return
        with capman.global_and_fixture_disabled():
            tr.write_line(line)
            tr.flush()

    return say


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
