"""Import helpers: the package directory ``cloudrendering-vulkan_amd`` is not a
valid Python identifier, so it is loaded by path under the name ``crmi355x``.
"""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cloudrendering-vulkan_amd")
PKG_NAME = "crmi355x"


def package():
    if PKG_NAME in sys.modules:
        return sys.modules[PKG_NAME]
    spec = importlib.util.spec_from_file_location(PKG_NAME, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[PKG_NAME] = mod
    spec.loader.exec_module(mod)
    return mod


def configs():
    package()
    import importlib
    return importlib.import_module(PKG_NAME + ".configs")


def oracle(lib_path=None):
    """The CPU oracle (test infrastructure): only tests/, smoke() and the
    bench's cpu_baseline leg call this.  lib_path: another build of the same
    sources (the bench's -march=native build)."""
    pkg = package()
    spec = importlib.util.spec_from_file_location("cr_oracle_py", os.path.join(ROOT, "oracle", "oracle_py.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.Oracle(pkg.structs, lib_path)
