// k_spmm.hip -- block SpMM with the Laplacian, the replacement of
// CsrOperator::apply / spmv_csr / sprs::prod::csr_mulacc_dense_rowmaj
// (reference: src/operator.rs:24-57, 176-228) and of the Chebyshev
// recurrence step (src/solvers/lobpcg.rs:113-150) through a fused epilogue:
//
//     out = s_l * (L x) + s_x * x + s_z * z,   L x = x - A x  (unit diagonal folded)
//
// Three schedules, picked per width and row length:
//   narrow (k <= 8)   G lanes per row, every lane gathers whole x rows;
//                     G-lane butterfly reduction; 64/G rows per wavefront.
//   wide (k >= 16)    16-byte chunks of the x row across C = k/VEC lanes, the
//                     remaining lanes split the row's nonzeros (slots); rows
//                     per wavefront = 64 / (C * slots); slot reduction by DPP.
//   long rows         rows with > kLongRow entries (degree hubs) get a 256-thread
//                     workgroup each, appended to the same launch, reduced by
//                     wave butterflies and a fixed-order sum of the 4 waves.
// Every reduction has a fixed order, so results are deterministic.
// Blocks are remapped so that each of the 8 XCDs streams a contiguous range
// of rows (its L2 then holds the x rows that neighbouring rows gather).  A
// launch gated by a device-planned round (the batched solver) uses only the
// resident workgroups, each walking its XCD's blocks, so that a launch the gate
// keeps idle costs a few microseconds instead of dispatching every block.
#include <hip/hip_ext.h>

#include <map>
#include <mutex>
#include <utility>

#include "si_host.hpp"

namespace si {

namespace {

SpmmProfile* g_profile = nullptr;  // set by the benchmark seam; host-side only

hipEvent_t pooled_event(SpmmProfile* p) {
    if (p->pool_next == p->pool.size()) {
        hipEvent_t e = nullptr;
        SI_HIP(hipEventCreate(&e));
        p->pool.push_back(e);
    }
    return p->pool[p->pool_next++];
}

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; static constexpr int n = 4; };
template <> struct Vec4<double> { using type = double2; static constexpr int n = 2; };

__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
    // Workgroups are dispatched round-robin over 8 XCDs; give XCD (b % 8) the
    // contiguous chunk [x*per, (x+1)*per) of the first 8*per blocks (a
    // bijection); the remaining nb % 8 blocks keep their own index.
    const int64_t per = nb / 8;
    if (b >= 8 * per) return b;
    return (b % 8) * per + b / 8;
}

// sign of the stored off-diagonal values: a32 holds +A (so L x = x - acc),
// l64 holds the negative L entries (so L x = x + acc).
template <typename T> __device__ inline T apply_diag(T xi, T acc);
template <> __device__ inline float apply_diag<float>(float xi, float acc) { return xi - acc; }
template <> __device__ inline double apply_diag<double>(double xi, double acc) { return xi + acc; }

// -------------------------------------------------------------- narrow
// Per-segment epilogue scalars: with rseg set, row i uses
// (s_l, s_x, s_z) = coef[3 * rseg[i] + {0, 1, 2}] instead of the arguments; a
// segment with s_l == 0 is a copy-through (no gathers).
template <typename T>
__device__ inline void seg_scalars(const int32_t* __restrict__ rseg, const T* __restrict__ coef,
                                   int64_t i, T& s_l, T& s_x, T& s_z) {
    if (rseg) {
        const int g = rseg[i];
        s_l = coef[3 * g];
        s_x = coef[3 * g + 1];
        s_z = coef[3 * g + 2];
    }
}

// Three-buffer ring of a device-planned filter: step j of a round whose gate
// (last step) is g reads x = the previous step's output (the round's input x0
// for j = 1), z = the output before that (x0 for j = 2) and writes
// ring[(g - j) mod 3], so the round's last step always writes ring[0].  On
// entry x is x0 and z null.
template <typename T>
__device__ inline void ring_buffers(int g, int j, T* r0, T* r1, T* r2, const T*& x, T*& out,
                                    const T*& z) {
    auto pick = [&](int d) -> T* {
        const int i = ((d % 3) + 3) % 3;
        return i == 0 ? r0 : (i == 1 ? r1 : r2);
    };
    const T* x0 = x;
    out = pick(g - j);
    if (j >= 2) x = pick(g - j + 1);
    if (j == 2) z = x0;
    else if (j >= 3) z = pick(g - j + 2);
}

// -------------------------------------------------------------- long rows
// A row with more than kLongRow entries (a degree hub) gets a whole 256-thread
// workgroup, appended after the row blocks of the same launch: threads stride
// the row's entries, each wave reduces by butterfly, the 4 wave sums are added
// in order (deterministic).  16 columns per pass.
template <typename T>
__device__ void long_row(int64_t i, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                         const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ out,
                         const T* __restrict__ z, int ld, int k, T s_l, T s_x, T s_z,
                         const int32_t* __restrict__ rseg, const T* __restrict__ coef) {
    constexpr int CH = 16;
    __shared__ T red[4][CH];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int32_t b = rp[i], e = rp[i + 1];
    seg_scalars<T>(rseg, coef, i, s_l, s_x, s_z);
    if (s_l == T(0)) e = b;
    for (int c0 = 0; c0 < k; c0 += CH) {
        const int cw = (k - c0) < CH ? (k - c0) : CH;
        T acc[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = T(0);
        for (int32_t p = b + t; p < e; p += 256) {
            const T v = val[p];
            const T* xr = x + (int64_t)col[p] * ld + c0;
#pragma unroll
            for (int c = 0; c < CH; ++c)
                if (c < cw) acc[c] += v * xr[c];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int c = 0; c < CH; ++c) acc[c] += __shfl_xor(acc[c], off, 64);
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < CH; ++c) red[wave][c] = acc[c];
        __syncthreads();
        if (t < cw) {
            const T sum = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
            const int64_t o = i * ld + c0 + t;
            const T xi = x[o];
            T r = s_l * apply_diag<T>(xi, sum) + s_x * xi;
            if (z) r += s_z * z[o];
            out[o] = r;
        }
        __syncthreads();
    }
}

// One block of rows (blockDim.x / G of them) of the narrow schedule.
template <typename T, int K, int G>
__device__ __forceinline__ void narrow_rows(
    int64_t blk, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ out,
    const T* __restrict__ z, int ld, int64_t n, T s_l, T s_x, T s_z,
    const int32_t* __restrict__ rseg, const T* __restrict__ coef) {
    const int lane = threadIdx.x % G;
    const int64_t i = (blk * blockDim.x + threadIdx.x) / G;
    if (i >= n) return;
    int32_t b = rp[i], e = rp[i + 1];
    if (e - b > kLongRow) return;
    seg_scalars<T>(rseg, coef, i, s_l, s_x, s_z);
    if (s_l == T(0)) e = b;
    T acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = T(0);
    // U nonzeros in flight per lane: indices first, then the gathers
    constexpr int U = 4;
    for (int32_t base = b + lane; base < e; base += G * U) {
        T vv[U];
        int32_t cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t p = base + u * G;
            const bool in = p < e;
            vv[u] = in ? val[p] : T(0);
            cc[u] = in ? col[p] : -1;
        }
        T xs[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cc[u] >= 0) {
                const T* xr = x + (int64_t)cc[u] * ld;
#pragma unroll
                for (int k = 0; k < K; ++k) xs[u][k] = xr[k];
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) xs[u][k] = T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += vv[u] * xs[u][k];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += __shfl_xor(acc[k], off, G);
    // lanes share the epilogue: lane l writes columns l, l+G, ...
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k % G != lane) continue;
        const int64_t o = i * ld + k;
        const T xi = x[o];
        T r = s_l * apply_diag<T>(xi, acc[k]) + s_x * xi;
        if (z) r += s_z * z[o];
        out[o] = r;
    }
}


// One workgroup per row block (blocks remapped per XCD); workgroups from
// nb_main on take one hub row each.
template <typename T, int K, int G>
__global__ __launch_bounds__(256) void k_spmm_narrow(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ x, T* __restrict__ out, const T* __restrict__ z, int ld, int64_t n,
    T s_l, T s_x, T s_z, const int32_t* __restrict__ rseg, const T* __restrict__ coef,
    const int32_t* __restrict__ long_rows, int64_t nb_main) {
    if ((int64_t)blockIdx.x >= nb_main) {  // appended hub-row workgroups
        long_row<T>(long_rows[blockIdx.x - nb_main], rp, col, val, x, out, z, ld, K, s_l, s_x,
                    s_z, rseg, coef);
        return;
    }
    narrow_rows<T, K, G>(xcd_block(blockIdx.x, nb_main), rp, col, val, x, out, z, ld, n, s_l,
                         s_x, s_z, rseg, coef);
}

// The gated form (launches of a device-planned round): the launch does
// nothing when gate_step > *gate; with a ring (r0 set) its buffers come from
// ring_buffers.  LOOP = false: one workgroup per row block, as above.  LOOP =
// true: workgroups [0, grid_main) walk the nb_main row blocks with stride
// grid_main (a multiple of 8, so every block stays on the XCD that xcd_block
// assigns it) -- an idle launch then dispatches only grid_main workgroups, a
// busy one runs a few percent slower.  Workgroups from grid_main on take one
// hub row each.
template <typename T, int K, int G, bool LOOP>
__global__ __launch_bounds__(256) void k_spmm_narrow_gated(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* x_arg, T* out_arg, const T* z_arg, int ld, int64_t n,
    T s_l, T s_x, T s_z, const int32_t* __restrict__ rseg, const T* __restrict__ coef,
    const int32_t* __restrict__ long_rows, int64_t nb_main, int64_t grid_main,
    const int32_t* __restrict__ gate, int gate_step, T* r0, T* r1, T* r2) {
    const T* x = x_arg;
    T* out = out_arg;
    const T* z = z_arg;
    const int g = *gate;
    if (gate_step > g) return;  // step past the round's filter degree
    if (r0 != nullptr) ring_buffers<T>(g, gate_step, r0, r1, r2, x, out, z);
    if ((int64_t)blockIdx.x >= grid_main) {  // appended hub-row workgroups
        long_row<T>(long_rows[blockIdx.x - grid_main], rp, col, val, x, out, z, ld, K, s_l, s_x,
                    s_z, rseg, coef);
        return;
    }
    if constexpr (LOOP) {
        for (int64_t bb = blockIdx.x; bb < nb_main; bb += grid_main)
            narrow_rows<T, K, G>(xcd_block(bb, nb_main), rp, col, val, x, out, z, ld, n, s_l,
                                 s_x, s_z, rseg, coef);
    } else {
        narrow_rows<T, K, G>(xcd_block(blockIdx.x, nb_main), rp, col, val, x, out, z, ld, n,
                             s_l, s_x, s_z, rseg, coef);
    }
}

// -------------------------------------------------------------- wide
// One block of rows (4 wavefronts x RPW rows) of the wide schedule.
template <typename T, int K, int SLOTS, int U>
__device__ __forceinline__ void wide_rows(
    int64_t blk, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ out,
    const T* __restrict__ z, int ld, int64_t n, T s_l, T s_x, T s_z,
    const int32_t* __restrict__ rseg, const T* __restrict__ coef) {
    using V = typename Vec4<T>::type;
    constexpr int VEC = Vec4<T>::n;
    constexpr int C = K / VEC;                 // lanes across one x row
    constexpr int RPW = 64 / (C * SLOTS);      // rows per wavefront
    static_assert(C * SLOTS * RPW == 64, "bad wide SpMM shape");
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = lane % C;
    const int slot = (lane / C) % SLOTS;
    const int rsub = lane / (C * SLOTS);
    const int64_t i = (blk * (blockDim.x >> 6) + wave) * RPW + rsub;
    const bool valid = i < n;
    int32_t b = 0, e = 0;
    if (valid) { b = rp[i]; e = rp[i + 1]; }
    const bool is_long = (e - b) > kLongRow;
    if (valid) seg_scalars<T>(rseg, coef, i, s_l, s_x, s_z);
    if (!valid || is_long || s_l == T(0)) { b = 0; e = 0; }
    T acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = T(0);
    // Memory-level parallelism: each lane first issues U (value, column) loads,
    // then U independent x-row gathers, then the FMAs -- one dependent
    // index->gather round trip per U*SLOTS nonzeros of the row.
    for (int32_t base = b + slot; base < e; base += SLOTS * U) {
        T vv[U];
        int32_t cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t p = base + u * SLOTS;
            const bool in = p < e;
            vv[u] = in ? val[p] : T(0);
            cc[u] = in ? col[p] : -1;
        }
        V xs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cc[u] >= 0) {
                xs[u] = *reinterpret_cast<const V*>(x + (int64_t)cc[u] * ld + chunk * VEC);
            } else {
                T* zr = reinterpret_cast<T*>(&xs[u]);
#pragma unroll
                for (int v = 0; v < VEC; ++v) zr[v] = T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T* a = reinterpret_cast<const T*>(&xs[u]);
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] += vv[u] * a[v];
        }
    }
    // reduce over slots (lane bits between C and C*SLOTS)
#pragma unroll
    for (int off = C; off < C * SLOTS; off <<= 1)
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[v] += __shfl_xor(acc[v], off, 64);
    if (slot == 0 && valid && !is_long) {
        const int64_t o = i * ld + chunk * VEC;
        const V xv = *reinterpret_cast<const V*>(x + o);
        const T* xi = reinterpret_cast<const T*>(&xv);
        V r;
        T* rr = reinterpret_cast<T*>(&r);
        if (z) {
            const V zv = *reinterpret_cast<const V*>(z + o);
            const T* zi = reinterpret_cast<const T*>(&zv);
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                rr[v] = s_l * apply_diag<T>(xi[v], acc[v]) + s_x * xi[v] + s_z * zi[v];
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) rr[v] = s_l * apply_diag<T>(xi[v], acc[v]) + s_x * xi[v];
        }
        *reinterpret_cast<V*>(out + o) = r;
    }
}


// Grid as k_spmm_narrow's.
template <typename T, int K, int SLOTS, int U>
__global__ __launch_bounds__(256) void k_spmm_wide(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ x, T* __restrict__ out, const T* __restrict__ z, int ld, int64_t n,
    T s_l, T s_x, T s_z, const int32_t* __restrict__ rseg, const T* __restrict__ coef,
    const int32_t* __restrict__ long_rows, int64_t nb_main) {
    if ((int64_t)blockIdx.x >= nb_main) {  // appended hub-row workgroups
        long_row<T>(long_rows[blockIdx.x - nb_main], rp, col, val, x, out, z, ld, K, s_l, s_x,
                    s_z, rseg, coef);
        return;
    }
    wide_rows<T, K, SLOTS, U>(xcd_block(blockIdx.x, nb_main), rp, col, val, x, out, z, ld, n,
                              s_l, s_x, s_z, rseg, coef);
}

// Gated form, as k_spmm_narrow_gated.
template <typename T, int K, int SLOTS, int U, bool LOOP>
__global__ __launch_bounds__(256) void k_spmm_wide_gated(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* x_arg, T* out_arg, const T* z_arg, int ld, int64_t n,
    T s_l, T s_x, T s_z, const int32_t* __restrict__ rseg, const T* __restrict__ coef,
    const int32_t* __restrict__ long_rows, int64_t nb_main, int64_t grid_main,
    const int32_t* __restrict__ gate, int gate_step, T* r0, T* r1, T* r2) {
    const T* x = x_arg;
    T* out = out_arg;
    const T* z = z_arg;
    const int g = *gate;
    if (gate_step > g) return;  // step past the round's filter degree
    if (r0 != nullptr) ring_buffers<T>(g, gate_step, r0, r1, r2, x, out, z);
    if ((int64_t)blockIdx.x >= grid_main) {  // appended hub-row workgroups
        long_row<T>(long_rows[blockIdx.x - grid_main], rp, col, val, x, out, z, ld, K, s_l, s_x,
                    s_z, rseg, coef);
        return;
    }
    if constexpr (LOOP) {
        for (int64_t bb = blockIdx.x; bb < nb_main; bb += grid_main)
            wide_rows<T, K, SLOTS, U>(xcd_block(bb, nb_main), rp, col, val, x, out, z, ld, n,
                                      s_l, s_x, s_z, rseg, coef);
    } else {
        wide_rows<T, K, SLOTS, U>(xcd_block(blockIdx.x, nb_main), rp, col, val, x, out, z, ld, n,
                                  s_l, s_x, s_z, rseg, coef);
    }
}

template <typename T>
struct Epi {
    T s_l, s_x, s_z;
    const T* z;
    const int32_t* rseg;
    const T* coef;
    // device-side gate: the launch does nothing when gate_step > *gate
    const int32_t* gate = nullptr;
    int gate_step = 0;
    // ring mode (with a gate): buffers from ring_buffers, x is the round's x0
    T* ring[3] = {nullptr, nullptr, nullptr};
    // resident workgroups walking the row blocks (cheap when the gate idles it)
    bool resident = false;
    // profiled launches: events recorded by the kernel's own dispatch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Launches `kernel`; with profiling events set, through hipExtLaunchKernelGGL,
// whose dispatch packet carries the start and stop timestamps (no separate
// marker packets between the solver's launches).
template <typename F, typename... Args>
void launch(F kernel, dim3 grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, Args... args) {
    if (ev0)
        hipExtLaunchKernelGGL(kernel, grid, dim3(256), 0, s, ev0, ev1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, args...);
}

// Workgroups of `kernel` (256 threads) resident on the current device at
// once, rounded down to a multiple of 8 (one per XCD slot); looked up once per
// kernel and device.
template <typename F>
int64_t resident_workgroups(F kernel) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int64_t> known;
    int dev = 0;
    SI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dev);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
    int cus = 0, per_cu = 0;
    SI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
    int64_t g = (int64_t)cus * (per_cu > 0 ? per_cu : 1);
    g = g / 8 * 8;
    if (g < 8) g = 8;
    known.emplace(key, g);
    return g;
}

// Row-block grid of a launch: one workgroup per block, or -- for a launch
// that a device gate may keep idle, where the empty launch must cost little --
// the resident workgroups walking the blocks.
template <typename F>
int64_t main_grid(F kernel, int64_t blocks, bool resident) {
    if (!resident) return blocks;
    const int64_t g = resident_workgroups(kernel);
    return blocks < g ? blocks : g;
}

template <typename T, int K, int G>
void launch_narrow(const DevLaplacian& L, const T* val, const T* x, T* out, int ld,
                   const Epi<T>& ep, hipStream_t s) {
    const int64_t threads = L.n * G;
    const int64_t blocks = ceil_div(threads, 256);
    if (ep.gate == nullptr) {
        launch(k_spmm_narrow<T, K, G>, dim3(blocks + L.n_long), s, ep.ev0, ep.ev1,
               (const int32_t*)L.rowptr, (const int32_t*)L.col, val, x, out, ep.z, ld, L.n,
               ep.s_l, ep.s_x, ep.s_z, ep.rseg, ep.coef, (const int32_t*)L.long_rows, blocks);
        return;
    }
    auto kern = ep.resident ? k_spmm_narrow_gated<T, K, G, true>
                            : k_spmm_narrow_gated<T, K, G, false>;
    const int64_t grid = main_grid(kern, blocks, ep.resident);
    launch(kern, dim3(grid + L.n_long), s, ep.ev0, ep.ev1,
           (const int32_t*)L.rowptr, (const int32_t*)L.col, val, x, out, ep.z, ld, L.n, ep.s_l,
           ep.s_x, ep.s_z, ep.rseg, ep.coef, (const int32_t*)L.long_rows, blocks, grid, ep.gate,
           ep.gate_step, ep.ring[0], ep.ring[1], ep.ring[2]);
}

template <typename T, int K, int SLOTS, int U>
void launch_wide(const DevLaplacian& L, const T* val, const T* x, T* out, int ld,
                 const Epi<T>& ep, hipStream_t s) {
    constexpr int VEC = Vec4<T>::n;
    constexpr int C = K / VEC;
    constexpr int RPW = 64 / (C * SLOTS);
    const int64_t waves = ceil_div(L.n, RPW);
    const int64_t blocks = ceil_div(waves, 4);
    if (ep.gate == nullptr) {
        launch(k_spmm_wide<T, K, SLOTS, U>, dim3(blocks + L.n_long), s, ep.ev0, ep.ev1,
               (const int32_t*)L.rowptr, (const int32_t*)L.col, val, x, out, ep.z, ld, L.n,
               ep.s_l, ep.s_x, ep.s_z, ep.rseg, ep.coef, (const int32_t*)L.long_rows, blocks);
        return;
    }
    auto kern = ep.resident ? k_spmm_wide_gated<T, K, SLOTS, U, true>
                            : k_spmm_wide_gated<T, K, SLOTS, U, false>;
    const int64_t grid = main_grid(kern, blocks, ep.resident);
    launch(kern, dim3(grid + L.n_long), s, ep.ev0, ep.ev1,
           (const int32_t*)L.rowptr, (const int32_t*)L.col, val, x, out, ep.z, ld, L.n, ep.s_l,
           ep.s_x, ep.s_z, ep.rseg, ep.coef, (const int32_t*)L.long_rows, blocks, grid, ep.gate,
           ep.gate_step, ep.ring[0], ep.ring[1], ep.ring[2]);
}

// One launch of a directly supported width (1, 2, 3, 4, 8, 16, 32, 64).
template <typename T>
void spmm_chunk(const DevLaplacian& L, const T* val, const T* x, T* out, int k, int ld,
                const Epi<T>& ep, hipStream_t s) {
    constexpr bool F32 = sizeof(T) == 4;
    switch (k) {
        case 1: launch_narrow<T, 1, 8>(L, val, x, out, ld, ep, s); break;
        case 2: launch_narrow<T, 2, 8>(L, val, x, out, ld, ep, s); break;
        case 3: launch_narrow<T, 3, 8>(L, val, x, out, ld, ep, s); break;
        case 4: launch_narrow<T, 4, 8>(L, val, x, out, ld, ep, s); break;
        // 8 columns stay on the narrow schedule in FP64 too: the wide (8, 4)
        // shape over 16-byte chunks measured 622 us against 539 us on the 1M graph
        case 8: launch_narrow<T, 8, 8>(L, val, x, out, ld, ep, s); break;
        // (SLOTS, U): SLOTS*U nonzero positions per row per pass; ~32 fits the
        // 15-30 entries of k-NN rows with few idle lanes (measured on MI355X)
        // FP32 16/32 columns: the shape chosen for this graph by spmm_autotune
        case 16:
            if constexpr (F32) {
                switch (L.shape16) {
                    case 1: launch_wide<T, 16, 8, 4>(L, val, x, out, ld, ep, s); break;   // RPW=2
                    case 2: launch_wide<T, 16, 4, 8>(L, val, x, out, ld, ep, s); break;   // RPW=4
                    case 3: launch_wide<T, 16, 2, 8>(L, val, x, out, ld, ep, s); break;   // RPW=8
                    default: launch_wide<T, 16, 4, 4>(L, val, x, out, ld, ep, s); break;  // C=4, RPW=4
                }
            } else {
                launch_wide<T, 16, 4, 8>(L, val, x, out, ld, ep, s);       // C=8, RPW=2
            }
            break;
        case 32:
            if constexpr (F32) {
                switch (L.shape32) {
                    case 1: launch_wide<T, 32, 2, 8>(L, val, x, out, ld, ep, s); break;   // RPW=4
                    case 2: launch_wide<T, 32, 4, 8>(L, val, x, out, ld, ep, s); break;   // RPW=2
                    case 3: launch_wide<T, 32, 8, 4>(L, val, x, out, ld, ep, s); break;   // RPW=1
                    default: launch_wide<T, 32, 4, 4>(L, val, x, out, ld, ep, s); break;  // C=8, RPW=2
                }
            } else {
                launch_wide<T, 32, 4, 8>(L, val, x, out, ld, ep, s);       // C=16, RPW=1
            }
            break;
        case 64:
            if constexpr (F32) launch_wide<T, 64, 4, 8>(L, val, x, out, ld, ep, s);   // C=16, RPW=1
            else launch_wide<T, 64, 2, 8>(L, val, x, out, ld, ep, s);       // C=32, RPW=1
            break;
        default:
            throw Error(kErrInternal, "spmm: no kernel for width " + std::to_string(k));
    }
    SI_LAUNCH_CHECK();
}

// Any width: split the columns into directly supported chunks, widest first
// (the 16-byte vector paths need 16-byte aligned chunk offsets and leading
// dimension; otherwise chunks of at most 8 columns are used).  Hub rows ride
// along as extra workgroups of every chunk's launch.
template <typename T>
void spmm_impl(const DevLaplacian& L, const T* val, const T* x, T* out, int k, int ld, T s_l,
               T s_x, T s_z, const T* z, hipStream_t s, const int32_t* rseg = nullptr,
               const T* coef = nullptr, const int32_t* gate = nullptr, int gate_step = 0) {
    SI_CHECK(k >= 1 && ld >= k, kErrInvalidArgument, "spmm: need 1 <= k <= ld");
    SI_CHECK(x != out && (z == nullptr || z != out), kErrInvalidArgument,
             "spmm: output must not alias an input");
    const bool prof = g_profile && g_profile->enabled;
    const bool vec_ok = (ld % Vec4<T>::n) == 0 &&
                        (reinterpret_cast<uintptr_t>(x) % 16) == 0 &&
                        (reinterpret_cast<uintptr_t>(out) % 16) == 0 &&
                        (z == nullptr || (reinterpret_cast<uintptr_t>(z) % 16) == 0);
    const int widths_vec[] = {64, 32, 16, 8, 4, 3, 2, 1};
    const int widths_narrow[] = {8, 4, 3, 2, 1};
    int off = 0;
    while (off < k) {
        const int rem = k - off;
        int w = 1;
        if (vec_ok) {
            for (int c : widths_vec)
                if (c <= rem) { w = c; break; }
        } else {
            for (int c : widths_narrow)
                if (c <= rem) { w = c; break; }
        }
        Epi<T> ep{s_l, s_x, s_z, z ? z + off : nullptr, rseg, coef, gate, gate_step};
        if (prof) {
            ep.ev0 = pooled_event(g_profile);
            ep.ev1 = pooled_event(g_profile);
            // algorithmic bytes of this launch
            const double bytes = (double)L.nnz * (sizeof(T) + 4) + 4.0 * (L.n + 1) +
                                 (double)sizeof(T) * L.n * w * (z ? 3.0 : 2.0);
            g_profile->pending.emplace_back(ep.ev0, ep.ev1);
            g_profile->pending_bytes.emplace_back(bytes, sizeof(T) == 4 ? 1 : 0);
        }
        spmm_chunk<T>(L, val, x + off, out + off, w, ld, ep, s);
        off += w;
    }
}

// -------------------------------------------------------------- gather probe
// The SpMM's gather stream alone, to measure the rate the memory system gives
// it: a quad of lanes per stored entry p reads the 64-byte FP32 row c of a
// 16-column block (16 B per lane), 4 entries per quad in flight, grid-stride
// over the entries; each lane sums what it reads and writes one float4.
//   mode 0: c = col[p], the Laplacian's own column stream (no values, row
//           pointers or output rows);
//   mode 1: c = a hash of (p, col[p]) over all n rows (no locality, the same
//           footprint);
//   mode 2: the same hash over the first min(n, 32768) rows (a 2 MiB footprint
//           that every XCD's L2 holds).
// Every c is in [0, n); out holds gridDim.x * 256 float4.
__device__ inline uint32_t probe_hash(uint32_t v) {
    v ^= v >> 16; v *= 0x7feb352dU; v ^= v >> 15; v *= 0x846ca68bU; v ^= v >> 16;
    return v;
}

__global__ __launch_bounds__(256) void k_gather_probe(const int32_t* __restrict__ col, int64_t count,
                                                      int64_t n, int mode, const float* __restrict__ x,
                                                      float4* __restrict__ out) {
    constexpr int U = 4;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int q = threadIdx.x & 3;
    const int64_t g = t >> 2;
    const int64_t stride = ((int64_t)gridDim.x * 256) >> 2;
    const uint32_t span = mode == 2 ? (uint32_t)(n < 32768 ? n : 32768) : (uint32_t)n;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t p0 = g; p0 < count; p0 += stride * U) {
        int64_t cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = p0 + u * stride;
            int64_t c = -1;
            if (p < count) {
                c = col[p];
                if (mode != 0) c = probe_hash((uint32_t)p * 2654435761u ^ (uint32_t)c) % span;
            }
            cc[u] = c;
        }
        float4 xs[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            xs[u] = cc[u] >= 0 ? *reinterpret_cast<const float4*>(x + cc[u] * 16 + q * 4)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc.x += xs[u].x; acc.y += xs[u].y; acc.z += xs[u].z; acc.w += xs[u].w;
        }
    }
    out[t] = acc;
}

// y[i] = sum_k data[k] x[indices[k]] over row i's entries in storage order,
// each product rounded, then added (contraction into fused multiply-adds is
// switched off for this kernel): the arithmetic of
// the reference's scalar spmv_csr (src/operator.rs:24-57), bit for bit.  One
// thread per row, grid stride.
__global__ __launch_bounds__(256) void k_csr_spmv_exact(int64_t n, const int64_t* __restrict__ ip,
                                                        const int32_t* __restrict__ ix,
                                                        const double* __restrict__ data,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        double acc = 0.0;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            const double prod = data[k] * x[ix[k]];
            acc = acc + prod;
        }
        y[i] = acc;
    }
}

}  // namespace

void csr_spmv_exact_f64(int64_t n, const int64_t* ip, const int32_t* ix, const double* data,
                        const double* x, double* y, hipStream_t s) {
    int64_t blocks = ceil_div(n, (int64_t)256);
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_csr_spmv_exact, dim3((unsigned)blocks), dim3(256), 0, s, n, ip, ix, data,
                       x, y);
    SI_LAUNCH_CHECK();
}

double gather_probe_us(const DevLaplacian& L, int mode, int reps, hipStream_t s) {
    SI_CHECK(mode >= 0 && mode <= 2, kErrInvalidArgument, "gather probe: mode 0, 1 or 2");
    SI_CHECK(reps >= 1, kErrInvalidArgument, "gather probe: reps >= 1");
    SI_CHECK(L.n > 0 && L.nnz > 0, kErrInvalidArgument, "gather probe: empty Laplacian");
    const unsigned grid = 1024;
    float* x = nullptr;
    float4* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double us = 0.0;
    try {
        SI_HIP(hipMalloc(&x, sizeof(float) * 16 * (size_t)(L.n > 0 ? L.n : 1)));
        SI_HIP(hipMalloc(&out, sizeof(float4) * grid * 256));
        SI_HIP(hipMemsetAsync(x, 0, sizeof(float) * 16 * (size_t)L.n, s));
        SI_HIP(hipEventCreate(&e0));
        SI_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, s, (const int32_t*)L.col, L.nnz,
                           L.n, mode, (const float*)x, out);  // warm
        SI_LAUNCH_CHECK();
        SI_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, s, (const int32_t*)L.col,
                               L.nnz, L.n, mode, (const float*)x, out);
        SI_LAUNCH_CHECK();
        SI_HIP(hipEventRecord(e1, s));
        SI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SI_HIP(hipEventElapsedTime(&ms, e0, e1));
        us = 1e3 * ms / reps;
    } catch (...) {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (x) hipFree(x);
        if (out) hipFree(out);
        throw;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(x);
    hipFree(out);
    return us;
}

void spmm_profile_attach(SpmmProfile* p) { g_profile = p; }

bool spmm_profiling() { return g_profile && g_profile->enabled; }

void spmm_profile_collect(SpmmProfile* p) {
    if (!p) return;
    for (size_t i = 0; i < p->pending.size(); ++i) {
        SI_HIP(hipEventSynchronize(p->pending[i].second));
        float ms = 0.f;
        SI_HIP(hipEventElapsedTime(&ms, p->pending[i].first, p->pending[i].second));
        p->ms += ms;
        p->bytes += p->pending_bytes[i].first;
        p->launches += 1;
        if (p->pending_bytes[i].second) {
            p->ms_f32 += ms;
            p->bytes_f32 += p->pending_bytes[i].first;
            p->launches_f32 += 1;
        }
    }
    p->pending.clear();
    p->pending_bytes.clear();
    p->pool_next = 0;  // every pooled event has completed: reuse them
}

void spmm_profile_release(SpmmProfile* p) {
    if (!p) return;
    for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
    p->pool.clear();
    p->pool_next = 0;
}

bool spmm_width_supported(int k) { return k >= 1 && k <= 4096; }

int spmm_shape_count() { return 4; }

void spmm_autotune(DevLaplacian& L, hipStream_t s, int widths) {
    const int64_t n = L.n;
    const int reps = 2;
    const int wmax = (widths & kTuned32) ? 32 : 16;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SpmmProfile* saved = g_profile;
    g_profile = nullptr;  // tuning launches are not part of a profiled region
    try {
        // the solvers' FP32 block buffers (grow-only workspace), refilled by
        // every solve
        float* x = ws_get<float>(kWsFastX32, (size_t)n * wmax);
        float* y = ws_get<float>(kWsFastAx32, (size_t)n * wmax);
        // any finite block: the x row values do not change the access pattern
        SI_HIP(hipMemsetAsync(x, 0, sizeof(float) * n * wmax, s));
        SI_HIP(hipEventCreate(&e0));
        SI_HIP(hipEventCreate(&e1));
        for (int width : {16, 32}) {
            if (!(widths & (width == 16 ? kTuned16 : kTuned32))) continue;
            float best = 0.f;
            int pick = 0;
            for (int c = 0; c < spmm_shape_count(); ++c) {
                (width == 16 ? L.shape16 : L.shape32) = c;
                spmm_f32(L, x, y, width, width, 1.f, 0.f, 0.f, nullptr, s);  // warm
                SI_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < reps; ++r) spmm_f32(L, x, y, width, width, 1.f, 0.f, 0.f, nullptr, s);
                SI_HIP(hipEventRecord(e1, s));
                SI_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                SI_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (c == 0 || ms < best) { best = ms; pick = c; }
            }
            (width == 16 ? L.shape16 : L.shape32) = pick;
            L.tuned |= width == 16 ? kTuned16 : kTuned32;
            if (trace_on())
                std::fprintf(stderr, "[autotune] FP32 width %d: shape %d (%.1f us/launch)\n", width,
                             pick, 1e3 * best / reps);
        }
    } catch (...) {
        g_profile = saved;
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (widths & kTuned16) L.shape16 = 0;
        if (widths & kTuned32) L.shape32 = 0;
        throw;
    }
    g_profile = saved;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
}

void spmm_f32(const DevLaplacian& L, const float* x, float* out, int k, int ld, float s_l,
              float s_x, float s_z, const float* z, hipStream_t s, const int32_t* gate,
              int gate_step) {
    spmm_impl<float>(L, L.a32, x, out, k, ld, s_l, s_x, s_z, z, s, nullptr, nullptr, gate,
                     gate_step);
}

void spmm_f64(const DevLaplacian& L, const double* x, double* out, int k, int ld, double s_l,
              double s_x, double s_z, const double* z, hipStream_t s, const int32_t* gate,
              int gate_step) {
    spmm_impl<double>(L, L.l64, x, out, k, ld, s_l, s_x, s_z, z, s, nullptr, nullptr, gate,
                      gate_step);
}

namespace {

template <typename T>
void spmm_ring_impl(const DevLaplacian& L, const T* val, const T* x0, T* const ring[3], int k,
                    int ld, const int32_t* rseg, const T* coef, const int32_t* gate, int step,
                    bool resident, hipStream_t s) {
    SI_CHECK(rseg && coef && gate && ring[0] && ring[1] && ring[2] && x0 && step >= 1,
             kErrInvalidArgument, "spmm ring step: buffers, segment scalars and gate required");
    SI_CHECK(ring[0] != ring[1] && ring[1] != ring[2] && ring[0] != ring[2] && x0 != ring[0] &&
                 x0 != ring[1] && x0 != ring[2],
             kErrInvalidArgument, "spmm ring step: the ring and x0 must be four distinct buffers");
    bool aligned = (ld % Vec4<T>::n) == 0 && (reinterpret_cast<uintptr_t>(x0) % 16) == 0;
    for (int i = 0; i < 3; ++i) aligned = aligned && (reinterpret_cast<uintptr_t>(ring[i]) % 16) == 0;
    const bool single = k == 1 || k == 2 || k == 3 || k == 4 || k == 8 ||
                        ((k == 16 || k == 32 || k == 64) && aligned);
    SI_CHECK(single && ld >= k, kErrInvalidArgument, "spmm ring step: width must be one launch");
    Epi<T> ep{T(1), T(0), T(0), nullptr, rseg, coef, gate, step};
    for (int i = 0; i < 3; ++i) ep.ring[i] = ring[i];
    ep.resident = resident;
    spmm_chunk<T>(L, val, x0, ring[0], k, ld, ep, s);
}

}  // namespace

void spmm_seg_ring_f32(const DevLaplacian& L, const float* x0, float* const ring[3], int k, int ld,
                       const int32_t* rseg, const float* coef, const int32_t* gate, int step,
                       bool resident, hipStream_t s) {
    spmm_ring_impl<float>(L, L.a32, x0, ring, k, ld, rseg, coef, gate, step, resident, s);
}

void spmm_seg_ring_f64(const DevLaplacian& L, const double* x0, double* const ring[3], int k,
                       int ld, const int32_t* rseg, const double* coef, const int32_t* gate,
                       int step, bool resident, hipStream_t s) {
    spmm_ring_impl<double>(L, L.l64, x0, ring, k, ld, rseg, coef, gate, step, resident, s);
}

void spmm_seg_f32(const DevLaplacian& L, const float* x, float* out, int k, int ld,
                  const float* z, const int32_t* rseg, const float* coef, hipStream_t s,
                  const int32_t* gate, int gate_step) {
    SI_CHECK(rseg && coef, kErrInvalidArgument, "spmm_seg: segment map and scalars required");
    spmm_impl<float>(L, L.a32, x, out, k, ld, 1.f, 0.f, 0.f, z, s, rseg, coef, gate, gate_step);
}

void spmm_seg_f64(const DevLaplacian& L, const double* x, double* out, int k, int ld,
                  const double* z, const int32_t* rseg, const double* coef, hipStream_t s,
                  const int32_t* gate, int gate_step) {
    SI_CHECK(rseg && coef, kErrInvalidArgument, "spmm_seg: segment map and scalars required");
    spmm_impl<double>(L, L.l64, x, out, k, ld, 1.0, 0.0, 0.0, z, s, rseg, coef, gate, gate_step);
}

}  // namespace si
