import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as _ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C ABI")
    config.addinivalue_line("markers", "slow: full-size (64^3) scene")


@pytest.fixture(scope="session")
def pkg():
    return _ge.load_package()


@pytest.fixture(scope="session")
def orc():
    return _ge.load_oracle()


@pytest.fixture(scope="session")
def scene_tiny(pkg):
    return pkg.scenes.scene_tiny(N=6, n_pairs=12)


@pytest.fixture(scope="session")
def scene_small(pkg):
    """8^3 grid with stiffness spanning 1e4..1e9 and 40 contacts (several MAS levels at B=16)."""
    return pkg.scenes.tet_grid_scene("small_grid8", 8, E_loguniform=(1e4, 1e9), n_pairs=40)


@pytest.fixture(scope="session")
def scene_c0(pkg):
    return pkg.scenes.scene_c0()


@pytest.fixture(scope="session")
def gpu_ctx(pkg):
    ctx = pkg.Context(0)
    yield ctx
    ctx.close()
