"""MI355X-native Monte Carlo engine for QIOToolkit's annealing hot path.

The Metropolis sweep runs as hand-written HIP for gfx950 in libqmc.so behind
the C ABI of include/qmc.h; this package holds the ctypes binding and the
host-side mirror of the reference's solver interface.
"""
