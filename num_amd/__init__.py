"""num_amd -- MI355X-native CSR SpMV, multi-RHS SpMV and CG.

``num_amd.kokkos_spmv`` is the drop-in for the reference's ``kokkos_spmv``
module; it calls the hand-written gfx950 kernels of ``libnum_amd.so`` through
the C ABI in ``include/num_amd.h``.
"""
__version__ = "0.1.0"
