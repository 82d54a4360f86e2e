"""Loads the package module nitroozk-stwo_amd/stwo_hip.py (the directory name is not a Python
identifier, so it is imported by path)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nitroozk-stwo_amd")


def module():
    spec = importlib.util.spec_from_file_location("stwo_hip", os.path.join(PKG, "stwo_hip.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


SH = module()
