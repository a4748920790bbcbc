import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def kidney():
    """Kidney-phantom fixtures (reference config/Data) with their manifest."""
    import json
    d = os.path.join(GOLDEN, "kidney")
    with open(os.path.join(d, "kidney.json")) as fh:
        man = json.load(fh)

    def load(name, dims):
        return np.fromfile(os.path.join(d, name), np.float32).reshape(dims[1], dims[0])

    out = {"manifest": man,
           "ct": load("CT_KidneyPhantom.raw", man["fixed"]["dims"]),
           "us": load("US_KidneyPhantom.raw", man["moving"]["dims"])}
    for name, sp in man["identity_dumps"].items():
        vd = (int(np.ceil(np.float32(man["fixed"]["dims"][0] * np.float32(man["fixed"]["spacing"][0])) / np.float32(sp))),
              int(np.ceil(np.float32(man["fixed"]["dims"][1] * np.float32(man["fixed"]["spacing"][1])) / np.float32(sp))))
        out[name] = load(name, vd)
    for name, rec in man["rigid_dumps"].items():
        out[name] = load(name, rec["virtual_dims"])
    return out


@pytest.fixture(scope="session")
def engine():
    from lc2registration_amd import engine as E
    e = E.Engine(0)
    yield e
    e.close()
