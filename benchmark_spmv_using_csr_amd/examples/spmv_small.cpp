// The reference's built-in example (spmv_cuda/main.cu:431-545, `./spmv 0`)
// through the C++ operator class: 6 x 6 CSR, x = 1, expected y = [6, 39, 21,
// 0, 12, 42].  Unlike the reference (which never calls get_y, main.cu:499-531)
// it reads y back, and runs the SpMV twice to show the result is repeatable.
// Build: g++ -std=c++17 -I include spmv_small.cpp -L <pkg> -lbhsparse_spmv
// (add -DUSE_SVM_ALWAYS=1 for the unified-memory build, as in spmv_cuda/Makefile:6).
#include <cstdio>

#include "bhsparse_spmv.hpp"

int main()
{
    int row_ptr[7] = {0, 3, 9, 11, 11, 12, 15};
    int col_idx[15] = {0, 2, 5, 0, 1, 2, 3, 4, 5, 2, 4, 4, 2, 3, 4};
    value_type val[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    value_type x[6] = {1, 1, 1, 1, 1, 1};
    value_type y[6] = {0, 0, 0, 0, 0, 0};
    const value_type expect[6] = {6, 39, 21, 0, 12, 42};

    // the reference driver's own lines (main.cu:490-498), unchanged
    bhsparse_spmv_cuda *bhsparse = new bhsparse_spmv_cuda();
    int err = bhsparse->init_platform();
    if (err) {
        std::printf("init_platform failed: %d (%s)\n", err, bhs_last_error());
        delete bhsparse;
        return 1;
    }
    err = bhsparse->prepare_mem(6, 6, 15, row_ptr, col_idx, val, x, y);
    int bad = 0;
    for (int call = 0; call < 2 && !err; ++call) {
        err = bhsparse->run_benchmark();
        if (!err) err = bhsparse->get_y();
        for (int i = 0; i < 6; ++i) bad += y[i] != expect[i];
    }
    int mode = -1;
    bhsparse->getMemoryMode(&mode);
    std::printf("memory: %s\n", mode == BHS_MEMORY_MANAGED       ? "managed"
                                 : mode == BHS_MEMORY_HOST_MAPPED ? "host-mapped"
                                                                  : "device");
    std::printf("y = [%g, %g, %g, %g, %g, %g] %s\n", (double)y[0], (double)y[1], (double)y[2],
                (double)y[3], (double)y[4], (double)y[5], (!err && !bad) ? "PASS" : "FAIL");
    bhsparse->free_platform();
    bhsparse->free_mem();
    delete bhsparse;
    if (err || bad) return 1;

    // the OpenCL driver's sequence (spmv_opencl_amd/main.cpp:289-361): init_kernels
    // after init_platform, get_y checked, sync_device after the timed calls
    bhsparse_spmv_opencl *ocl = new bhsparse_spmv_opencl();
    if (ocl->init_kernels("-DUSE_DOUBLE=1") != BHSPARSE_NOT_INITIALIZED) bad++;   // call order
    err = ocl->init_platform();
    if (!err) err = ocl->init_kernels("-DUSE_DOUBLE=1");
    for (int i = 0; i < 6; ++i) y[i] = -1;
    if (!err) err = ocl->prepare_mem(6, 6, 15, row_ptr, col_idx, val, x, y);
    if (!err) err = ocl->run_benchmark();
    if (!err) err = ocl->get_y();
    for (int i = 0; i < 6 && !err; ++i) bad += y[i] != expect[i];
    for (int r = 0; r < 10 && !err; ++r) err = ocl->run_benchmark();
    ocl->sync_device();
    std::printf("opencl sequence: y = [%g, %g, %g, %g, %g, %g] %s\n", (double)y[0], (double)y[1],
                (double)y[2], (double)y[3], (double)y[4], (double)y[5],
                (!err && !bad) ? "PASS" : "FAIL");
    ocl->free_platform();
    ocl->free_mem();
    delete ocl;
    return (err || bad) ? 1 : 0;
}
