"""MI355X-native spectral-element time stepping for SPECFEM3D Cartesian's single-partition
solver and adjoint loops (NGLL = 5, FP32), behind the reference's Fortran-callable C-ABI."""
__version__ = "0.1.0"
