"""MI355X-native FP64 steady residual + LU-SGS path of RocketCombustor.

Modules:
    deck      reference input decks (species/thermo/kinetics/inlet/solver settings)
    nozzle    seeded synthetic nozzle cases (mesh, BCs, initial field)
    native    ctypes binding of the HIP library (C-ABI, include/rocketcombustor.h)
    solver    host-side mirror of the reference's steady operators on one case
    build     in-tree build of the HIP library for gfx950
"""
__version__ = "0.1.0"
