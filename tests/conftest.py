"""Test configuration: repository root on sys.path, the `gpu` marker and shared fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture
def tiled_raster():
    """Every resolution class's main view through the tiled raster for the
    test (the tuning file leaves some classes on the per-meshlet raster);
    the classes' tunings are restored afterwards."""
    from lighthugger_amd.lib import get_tuning_classes, set_tuning_classes

    saved = get_tuning_classes()
    forced = get_tuning_classes()
    for t in forced:
        t.tile_raster = 1
    set_tuning_classes(forced)
    yield
    set_tuning_classes(saved)
