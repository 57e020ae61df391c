// si_abi.hpp -- definitions of the opaque C ABI handles.
#pragma once

#include "si_host.hpp"
#include "../../include/spectral_mi355x.h"

struct si_context {
    si::Context c;
};

struct si_laplacian {
    si::DevLaplacian L;
    int device = 0;
};

namespace si {
// Blocks cross the ABI in the caller's node order; inside, rows follow the
// Laplacian's locality order (L.perm[new] = old).
inline void rows_to_internal(const DevLaplacian& L, const double* src, int lds, double* dst,
                             int ldd, int k, hipStream_t s) {
    gather_rows_f64(src, lds, dst, ldd, L.n, k, L.perm, s);
}
inline void rows_to_external(const DevLaplacian& L, const double* src, int lds, double* dst,
                             int ldd, int k, hipStream_t s) {
    gather_rows_f64(src, lds, dst, ldd, L.n, k, L.iperm, s);
}
inline void rows_to_internal(const DevLaplacian& L, const float* src, int lds, float* dst,
                             int ldd, int k, hipStream_t s) {
    gather_rows_f32(src, lds, dst, ldd, L.n, k, L.perm, s);
}
inline void rows_to_external(const DevLaplacian& L, const float* src, int lds, float* dst,
                             int ldd, int k, hipStream_t s) {
    gather_rows_f32(src, lds, dst, ldd, L.n, k, L.iperm, s);
}
}  // namespace si
