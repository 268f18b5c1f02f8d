"""Builds the test USER-source plug-in (tests/user_source/ray_plugin.c)."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(dest_dir):
    """Compiles the plug-in into dest_dir; returns (path, loaded CDLL)."""
    out = os.path.join(dest_dir, "ray_plugin.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "user_source", "ray_plugin.c"), "-lm"], check=True)
    return out, ctypes.CDLL(out)
