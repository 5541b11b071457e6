"""MI355X-native coherent F-statistic search core (polgraw-allsky network job_core).

Submodules:
  abi    - ctypes mirror of include/pgw.h
  pgw    - the HIP library (libpgw_hip.so) through its C ABI
  synth  - seeded synthetic segments (noise, ephemeris, grid, injection)
  build  - compiles the HIP library for gfx950

The directory name is not a Python identifier; load it with ``load()`` from
``pgwload.py`` at the repository root, which registers it as ``polgraw_allsky_amd``.
"""
