"""MI355X-native PSpectCosmo lattice time step (HIP kernels behind a C ABI)."""
