"""Shared helpers for the tests: module loading and seeded frames."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "svt-jpeg-xs_amd")


def load_package():
    if "svt_jpeg_xs_amd" in sys.modules:
        return sys.modules["svt_jpeg_xs_amd"]
    spec = importlib.util.spec_from_file_location("svt_jpeg_xs_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["svt_jpeg_xs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_synth():
    load_package()
    import svt_jpeg_xs_amd.synth as synth
    return synth


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    return oracle_py
