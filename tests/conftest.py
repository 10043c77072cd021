import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "pic-deposition_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def hip():
    import pichip

    return pichip


_PARITY = []


@pytest.fixture(scope="session")
def parity_log():
    """Appends (case, metric, value) lines that the end of the run prints
    (pytest's terminal summary), so parity figures show in quiet runs too."""
    return _PARITY.append


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if _PARITY:
        terminalreporter.section("parity figures")
        for line in _PARITY:
            terminalreporter.write_line(line)
