"""Shared test setup: import paths, the `gpu` marker, small helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run on the MI355X box)")


def oracle_layers(layers):
    """synth.LayerModel list -> oracle Layer list."""
    import ub_oracle as O
    return [O.Layer(l.mode, tuple(float(v) for v in l.mean), tuple(float(v) for v in np.ravel(l.precision)),
                    tuple(l.comp)) for l in layers]


def layer_infos(layers):
    """synth.LayerModel list -> unblending_amd.LayerInfo list."""
    import unblending_amd as U
    return [U.LayerInfo(U.CompOp(*l.comp), U.BlendMode(l.mode), U.GaussianColorModel(l.mean, l.precision))
            for l in layers]


@pytest.fixture(scope="session")
def oracle():
    import ub_oracle
    ub_oracle.lib()
    return ub_oracle
