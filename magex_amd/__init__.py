"""magex_amd: magex's atomistic stochastic-LLG hot path on AMD Instinct MI355X (gfx950).

Drop-in modules for the magex API on this path:
    magex_amd.models     make_model, write_model, LATTICES          (magex.models)
    magex_amd.spinmodel  UnitCellModel, load_results, SpinSystem    (magex.spinmodel)
    magex_amd.terms      CubicAnisotropy, UniaxialHigher            (magex.terms, local terms)
    magex_amd.llg        LLGSolver, llg_rhs, save_spins             (magex.llg, Heun)
    magex_amd.texture    topological_charge                         (magex.texture)
Numerics run in libmxllg.so (hand-written HIP kernels behind the C ABI of include/mxllg.h).
"""
__version__ = "0.1.0"

from .constants import GAMMA, K_B, MU_B  # noqa: F401
