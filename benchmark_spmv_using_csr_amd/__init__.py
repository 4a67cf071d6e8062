"""MI355X (gfx950) speculative segmented-sum CSR SpMV, y = alpha * A * x.

Native code lives in ``libbhsparse_spmv.so`` (HIP kernels + the C ABI of
``include/bhsparse_spmv.h``) and ``libbhs_host.so`` (seeded generators); build
both with ``__graft_entry__.build()``.  ``spmv.bhsparse_spmv`` mirrors the
reference's operator class (spmv_cuda/bhsparse_spmv_cuda.h:7-55).
"""
__all__ = ["spmv", "matgen"]
