// k_mpnn.hip -- kernels of the learned warm-start network (message passing,
// hidden 64, tanh).  No reference counterpart: the reference starts LOBPCG
// from N(0,1) plus the sqrt-degree hint (src/solvers/lobpcg.rs:243-252); this
// network predicts the two non-trivial eigenvector directions instead.
//
// Dense products run on the bf16 matrix cores: v_mfma_f32_16x16x16_bf16 with
// FP32 accumulation.  Operand layout (measured on gfx950 with an integer
// probe): A[i][k] sits in lane i + 16*(k/4), element k%4; B[k][j] in lane
// j + 16*(k/4), element k%4; D[i][j] in lane j + 16*(i/4), register i%4.
#include "si_device.hpp"

namespace si {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> bf16 bits, round to nearest even (finite inputs)
__device__ inline short bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// out[n x Q] = epi(A[n x K] Wa[K x Q] + B[n x K] Wb[K x Q] + bias + C)
//   act 0: identity, 1: tanh, 2: multiply by (1 - G^2) (tanh derivative at G)
// A, B, C, G, out FP32 row-major (leading dimensions K / Q); Wa, Wb bf16
// [K x Q] row-major.  256 threads = 4 waves; a wave owns 16 rows, the
// workgroup 64 rows.
template <int K, int Q>
__global__ __launch_bounds__(256) void k_mpnn_gemm(int64_t n, const float* __restrict__ a,
                                                   const short* __restrict__ wa,
                                                   const float* __restrict__ b,
                                                   const short* __restrict__ wb,
                                                   const float* __restrict__ bias,
                                                   const float* __restrict__ c, int act,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ out) {
    static_assert(K % 16 == 0 && Q % 16 == 0 && K <= 64 && Q <= 64, "mpnn gemm shape");
    constexpr int QT = Q / 16;
    __shared__ short swa[K * Q];
    __shared__ short swb[K * Q];
    const int t = threadIdx.x;
    for (int e = t; e < K * Q; e += 256) {
        swa[e] = wa[e];
        swb[e] = wb ? wb[e] : (short)0;
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 64 + wave * 16;
    const int64_t row = r0 + (lane & 15);
    const bool rok = row < n;
    const int kq = 4 * (lane >> 4);
    v4f acc[QT];
#pragma unroll
    for (int q = 0; q < QT; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int term = 0; term < 2; ++term) {
        const float* src = term == 0 ? a : b;
        if (!src) break;
        const short* w = term == 0 ? swa : swb;
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks) {
            const int k0 = ks * 16 + kq;
            float4 av = rok ? *reinterpret_cast<const float4*>(src + row * K + k0)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
            v4s ao;
            ao[0] = bf16_bits(av.x); ao[1] = bf16_bits(av.y);
            ao[2] = bf16_bits(av.z); ao[3] = bf16_bits(av.w);
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                const int col = q * 16 + (lane & 15);
                v4s bo;
#pragma unroll
                for (int e = 0; e < 4; ++e) bo[e] = w[(k0 + e) * Q + col];
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ao, bo, acc[q], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < QT; ++q) {
        const int col = q * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t orow = r0 + 4 * (lane >> 4) + r;
            if (orow >= n) continue;
            const int64_t o = orow * Q + col;
            float v = acc[q][r];
            if (bias) v += bias[col];
            if (c) v += c[o];
            if (act == 1) v = tanhf(v);
            else if (act == 2) { const float gv = g[o]; v *= 1.f - gv * gv; }
            out[o] = v;
        }
    }
}

// y = x * (1 - h^2), elementwise
__global__ void k_mpnn_dtanh(int64_t count, const float* __restrict__ x, const float* __restrict__ h,
                             float* __restrict__ y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) y[e] = x[e] * (1.f - h[e] * h[e]);
}

// Row features 0-5 and 14-15 (the probes 6-13 are written by the host loop):
// log1p(deg), log1p(#neighbours), mean / max / min / std of the row's A_hat
// entries (a32), sqrt(deg) * inv_max_sq, 1.
__global__ void k_mpnn_rowfeat(const int32_t* __restrict__ rp, const float* __restrict__ a32,
                               const double* __restrict__ deg, int64_t n, double inv_max_sq,
                               float* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = rp[i], e = rp[i + 1];
    double sum = 0.0, mx = 0.0, mn = 0.0;
    for (int32_t p = b; p < e; ++p) {
        const double v = a32[p];
        sum += v;
        if (p == b || v > mx) mx = v;
        if (p == b || v < mn) mn = v;
    }
    const double cnt = (double)(e - b);
    const double mean = cnt > 0 ? sum / cnt : 0.0;
    double ss = 0.0;
    for (int32_t p = b; p < e; ++p) ss += (a32[p] - mean) * (a32[p] - mean);
    const double d = deg[i] > 0.0 ? deg[i] : 0.0;
    float* fi = f + i * 16;
    fi[0] = (float)log1p(d);
    fi[1] = (float)log1p(cnt);
    fi[2] = (float)mean;
    fi[3] = (float)mx;
    fi[4] = (float)mn;
    fi[5] = (float)(cnt > 0 ? sqrt(ss / cnt) : 0.0);
    fi[14] = (float)(sqrt(d) * inv_max_sq);
    fi[15] = 1.f;
}

// column j of f [n x 16] = scale * p
__global__ void k_mpnn_setcol(int64_t n, int j, const float* __restrict__ p, float scale,
                              float* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i * 16 + j] = p[i] * scale;
}

// p -= alpha * v  (FP64), then pf = (float) p
__global__ void k_mpnn_deflate(int64_t n, double* __restrict__ p, const double* __restrict__ v,
                               double alpha, float* __restrict__ pf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = p[i] - alpha * v[i];
    p[i] = r;
    pf[i] = (float)r;
}

// dX = [X | V] coef:  dx[i][j] = sum_t X[i][t] cx[t][j] + sum_t V[i][t] cv[t][j]
// dX = X P + V Q row by row: P [8 x 8], Q [2 x 8] (the coverage-loss gradient)
__global__ void k_mpnn_dx(int64_t n, const float* __restrict__ x, const double* __restrict__ v,
                          MpnnDxCoef c, float* __restrict__ dx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi[kMpnnOut];
#pragma unroll
    for (int p = 0; p < kMpnnOut; ++p) xi[p] = x[kMpnnOut * i + p];
    const double w0 = v[2 * i], w1 = v[2 * i + 1];
#pragma unroll
    for (int q = 0; q < kMpnnOut; ++q) {
        double d = w0 * c.q[q] + w1 * c.q[kMpnnOut + q];
#pragma unroll
        for (int p = 0; p < kMpnnOut; ++p) d += xi[p] * c.p[p * kMpnnOut + q];
        dx[kMpnnOut * i + q] = (float)d;
    }
}

// FP32 master weights -> bf16 [K x Q] (wt = 0) or its transpose [Q x K] (wt = 1)
__global__ void k_mpnn_pack(const float* __restrict__ w, int kk, int qq, int wt,
                            short* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kk * qq) return;
    const int k = e / qq, q = e % qq;
    const short v = bf16_bits(w[e]);
    if (wt) out[q * kk + k] = v;
    else out[e] = v;
}

// Adam (bias-corrected), FP32 state, FP64 gradients
__global__ void k_mpnn_adam(int64_t count, float* __restrict__ p, const double* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float gi = (float)g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline dim3 g256(int64_t count) { return dim3((unsigned)ceil_div(count, 256)); }

}  // namespace

void mpnn_gemm(int64_t n, int k, int q, const float* a, const short* wa, const float* b,
               const short* wb, const float* bias, const float* c, int act, const float* g,
               float* out, hipStream_t s) {
    const dim3 grid((unsigned)ceil_div(n, 64));
    if (k == 16 && q == 64)
        hipLaunchKernelGGL((k_mpnn_gemm<16, 64>), grid, dim3(256), 0, s, n, a, wa, b, wb, bias, c, act, g, out);
    else if (k == 64 && q == 64)
        hipLaunchKernelGGL((k_mpnn_gemm<64, 64>), grid, dim3(256), 0, s, n, a, wa, b, wb, bias, c, act, g, out);
    else if (k == 64 && q == 16)
        hipLaunchKernelGGL((k_mpnn_gemm<64, 16>), grid, dim3(256), 0, s, n, a, wa, b, wb, bias, c, act, g, out);
    else
        throw Error(kErrInternal, "mpnn_gemm: unsupported shape");
    SI_LAUNCH_CHECK();
}

void mpnn_dtanh(int64_t count, const float* x, const float* h, float* y, hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_dtanh, g256(count), dim3(256), 0, s, count, x, h, y);
    SI_LAUNCH_CHECK();
}

void mpnn_rowfeat(const DevLaplacian& L, double inv_max_sq, float* f, hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_rowfeat, g256(L.n), dim3(256), 0, s, L.rowptr, L.a32, L.deg, L.n,
                       inv_max_sq, f);
    SI_LAUNCH_CHECK();
}

void mpnn_setcol(int64_t n, int j, const float* p, float scale, float* f, hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_setcol, g256(n), dim3(256), 0, s, n, j, p, scale, f);
    SI_LAUNCH_CHECK();
}

void mpnn_deflate(int64_t n, double* p, const double* v, double alpha, float* pf, hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_deflate, g256(n), dim3(256), 0, s, n, p, v, alpha, pf);
    SI_LAUNCH_CHECK();
}

void mpnn_dx(int64_t n, const float* x, const double* v, const MpnnDxCoef& c, float* dx,
             hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_dx, g256(n), dim3(256), 0, s, n, x, v, c, dx);
    SI_LAUNCH_CHECK();
}

void mpnn_pack(const float* w, int k, int q, int transpose, short* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mpnn_pack, g256((int64_t)k * q), dim3(256), 0, s, w, k, q, transpose, out);
    SI_LAUNCH_CHECK();
}

void mpnn_adam(int64_t count, float* p, const double* g, float* m, float* v, float lr, float b1,
               float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_mpnn_adam, g256(count), dim3(256), 0, s, count, p, g, m, v, lr, b1, b2,
                       eps, c1, c2);
    SI_LAUNCH_CHECK();
}

}  // namespace si
