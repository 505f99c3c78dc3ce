import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "sunny.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libsunnyhip kernels)")
    config.addinivalue_line("markers", "slow: long CPU-only statistical or large-size checks")


import pytest  # noqa: E402

# Device layouts of the :dipole gather exercised by the parity tests:
# automatic plan, ghost-padded global gather, and forced LDS tile shapes
# (partial tiles, several axis-3 segments).  The forced shapes apply where
# they fit a model; elsewhere the library falls back to the ghost layout.
LAYOUTS = {
    "auto": ({}, 0),
    "ghost": ({}, 1),
    "tile2x2": ({"SNH_TILE_T1": "2", "SNH_TILE_T2": "2"}, 0),
    "tile4x2seg2": ({"SNH_TILE_T1": "4", "SNH_TILE_T2": "2", "SNH_TILE_NSEG": "2"}, 0),
}


@pytest.fixture(params=list(LAYOUTS))
def layout(request, monkeypatch):
    env, code = LAYOUTS[request.param]
    for k in ("SNH_TILE_T1", "SNH_TILE_T2", "SNH_TILE_NSEG"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    return code
