import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle_lib
    if not oracle_lib.ref_available():
        pytest.skip("reference build oracle/_ref is not present")
    return oracle_lib.ref()


@pytest.fixture(scope="session")
def engine():
    """The HIP engine; a GPU test fails (not skips) when it cannot load."""
    from xmcml_amd import engine as eng
    eng.lib()
    return eng
