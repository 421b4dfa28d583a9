import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rust-simulation-tools_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def cfg0():
    """2,500-atom synthetic complex (2,300 receptor + 200 ligand, 22 A sphere)
    and its first noisy frame (A)."""
    from rust_simulation_tools import synthetic
    top, base, rec, lig, seed = synthetic.config_system(0)
    return top, synthetic.noisy_frame(base, seed, 0), rec, lig, seed


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
