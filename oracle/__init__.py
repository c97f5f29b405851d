"""CPU oracle of the DFTTest/FFT3D core: test infrastructure only."""
