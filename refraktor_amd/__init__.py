"""refraktor_amd — MI355X-native fractal-flame renderer (librefrakt drop-in).

The native library (librefrakt_amd.so, C ABI in include/refrakt.h) does all the
work; this package holds the Python mirror of librefrakt's render API, the
seeded synthetic flame generator and the build script.
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CONFIG_DIR = PACKAGE_DIR / "config"
LIB_PATH = PACKAGE_DIR / "librefrakt_amd.so"

__all__ = ["PACKAGE_DIR", "CONFIG_DIR", "LIB_PATH"]
