"""CPU oracle of magex's LLG hot path: test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

magex_ref  NumPy/scipy restatement of magex's NumPy path (assembly, field, energy, Heun, charge)
philox     restatement of the counter-based noise generator the HIP kernels use
cpu        C++/OpenMP restatement of magex's numba fused Heun (cpu_llg.cpp), the CPU baseline
"""
