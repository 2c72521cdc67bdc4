import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def _ensure_built():
    from superdca_amd import build

    lib = os.path.join(ROOT, "superdca_amd", "libsdca.so")
    orc = os.path.join(ROOT, "oracle", "liboracle_plmdca.so")
    if not os.path.exists(orc):
        build.build_oracle()
    if not os.path.exists(lib):
        build.build_lib()


_ensure_built()


@pytest.fixture(scope="session")
def c1():
    """The C1 workload (B=200, N=1000) as envelope states."""
    import superdca_amd as sd
    from superdca_amd import synth

    d = synth.workload("C1")
    states, loci = sd.prepare_alignment(d["chars"])
    return {"chars": d["chars"], "states": states, "loci": loci, "pairs": d["pairs"]}


@pytest.fixture(scope="session")
def c4_states():
    """The genome-scale C4 workload (B=3000, N=94,000) as q_env = 5 states."""
    import superdca_amd as sd
    from superdca_amd import synth

    st, _ = sd.prepare_alignment(synth.workload("C4")["chars"], q=5)
    return st
