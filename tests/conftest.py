import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "hydro-bench_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


_CACHE = {}


def cached_mesh(name, dtype=np.float64, **kw):
    """Synthetic meshes are deterministic; build each once per session."""
    import hydro_amd

    key = (name, np.dtype(dtype).str, tuple(sorted(kw.items())))
    if key not in _CACHE:
        if name in hydro_amd.CONFIGS:
            _CACHE[key] = hydro_amd.make_config(name, dtype=dtype, **kw)
        elif name == "channel":
            _CACHE[key] = hydro_amd.make_channel_mesh(dtype=dtype, **kw)
        elif name == "floodplain":
            _CACHE[key] = hydro_amd.make_floodplain_mesh(dtype=dtype, **kw)
        else:
            raise KeyError(name)
    return _CACHE[key]


@pytest.fixture(scope="session")
def oracle():
    from oracle import hydro_oracle

    hydro_oracle.load()
    return hydro_oracle


def rel_err(got, ref):
    """Normwise relative difference max|got-ref| / max|ref| (0 for equal arrays)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.max(np.abs(got - ref)) if got.size else 0.0
    s = np.max(np.abs(ref)) if ref.size else 0.0
    if d == 0.0:
        return 0.0
    return d / s if s > 0 else np.inf
