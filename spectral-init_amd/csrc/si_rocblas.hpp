// si_rocblas.hpp -- the rocBLAS handle shared by the library's plain library
// calls (rocBLAS GEMM / TRSM, rocSOLVER eigen and Cholesky routines).
#pragma once

#include <rocblas/rocblas.h>

#include <string>

#include "si_common.hpp"

namespace si {

inline void rb_check(rocblas_status st, const char* what) {
    if (st != rocblas_status_success)
        throw Error(kErrHip, std::string(what) + ": rocBLAS status " + std::to_string((int)st));
}

// One rocBLAS handle per host thread (a context is used by one thread), kept
// until the process ends (destroying it in a thread-exit destructor could run
// after the HIP runtime has shut down); bound to stream s on every call.
inline rocblas_handle rb_handle(hipStream_t s) {
    static thread_local rocblas_handle h = nullptr;
    if (!h) rb_check(rocblas_create_handle(&h), "rocblas_create_handle");
    rb_check(rocblas_set_stream(h, s), "rocblas_set_stream");
    return h;
}

}  // namespace si
