"""openhomelbm_amd: the HOME-LBM time step (D2Q9 / D3Q27) on AMD MI355X.

Hand-written HIP kernels for gfx950 behind a C ABI (include/homelbm.h),
bound with ctypes; lattice state lives in PyTorch-ROCm tensors. The
reference-compatible API is in openhomelbm_amd.lbm (LBM_Solver, HomeFlow)
and openhomelbm_amd.lbm3d (LBM_Solver3D, HomeFlow3D).
"""

__version__ = "0.1.0"
