"""Shared test setup: the `gpu` marker, the package loader and seeded scenes."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "fuzzy-ellipses_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    """Import the package directory `fuzzy-ellipses_amd/` as `fuzzy_ellipses_amd`."""
    if "fuzzy_ellipses_amd" in sys.modules:
        return sys.modules["fuzzy_ellipses_amd"]
    spec = importlib.util.spec_from_file_location(
        "fuzzy_ellipses_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fuzzy_ellipses_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def in_record_order(scene, order):
    """The scene with its Gaussians in the renderer's record order of one view (record j is
    Gaussian order[j], from Renderer.debug_state); the order must be a permutation."""
    import dataclasses
    import numpy as np
    assert np.array_equal(np.sort(order), np.arange(scene.n)), "record order is not a permutation"
    return dataclasses.replace(scene, means=scene.means[order], prec=scene.prec[order],
                               weights_log=scene.weights_log[order])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def fea():
    return load_package()


@pytest.fixture(scope="session")
def scenes_mod(fea):
    import fuzzy_ellipses_amd.scenes as sc
    return sc
