// k_embed.hip -- the output stage of spectral_init on the device: eigenvector
// selection and sign fixing, the placement of every connected component, and
// the final scaling and noise.
//
//   select_eigenvectors + normalize_signs (src/selection.rs:11-29,
//   src/scaling.rs:62-74): the columns of each component's block in ascending
//   eigenvalue order (skipping the trivial one), each negated when its first
//   entry of largest magnitude, in node order, is negative.  The largest
//   magnitude is a maximum of the bit patterns of |x| (non-negative doubles
//   order as their bits), the first node attaining it a minimum of node ids.
//   embed_single_component (src/multi_component.rs:261-312): coordinates
//   sel * (range_c / max|sel|) + meta_c, or meta_c alone for components too
//   small for their own embedding.
//   scale_and_add_noise (src/scaling.rs:78-89): coordinates * (10 / max|c|),
//   rounded to f32, plus f32(1e-4 * N(0, 1)).
//
// The eigenvector block may be in any row order (the solver's locality order
// for the batched components): each component is a contiguous row range of
// it, split into work items of at most kChunk rows, so every maximum is a
// workgroup reduction plus one atomic per work item and column.  Every value
// is formed with the host code's operations in the same order (contraction
// off), so the output is bit-identical to it.  Non-finite eigenvector entries
// are left out of the maxima.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "si_host.hpp"

namespace si {

namespace {

constexpr int kChunk = 4096;     // rows per work item
constexpr int kMaxBlocks = 512;  // grid of the coordinate pass

__device__ __forceinline__ unsigned long long abs_bits(double x) {
    return (unsigned long long)(__double_as_longlong(x) & 0x7fffffffffffffffll);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffull), o);
        const unsigned hi = __shfl_xor((unsigned)(v >> 32), o);
        const unsigned long long w = ((unsigned long long)hi << 32) | lo;
        v = w > v ? w : v;
    }
    return v;
}

// maximum over the 256 threads of a workgroup (result valid in thread 0)
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v,
                                                            unsigned long long* red) {
    v = wave_max_u64(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) v = red[w] > v ? red[w] : v;
    }
    __syncthreads();
    return v;
}

struct Item {
    int64_t r0, r1;  // rows [r0, r1) of the block
    int32_t seg;
    int32_t pad;
};

// per (segment, column): bits of max |x| over the segment's rows
__global__ __launch_bounds__(256) void k_emb_colmax(const Item* __restrict__ items,
                                                    const double* __restrict__ b, int ldb,
                                                    const int32_t* __restrict__ cols, int d,
                                                    unsigned long long* __restrict__ cmax) {
    __shared__ unsigned long long red[4];
    const Item it = items[blockIdx.x];
    for (int j = 0; j < d; ++j) {
        const int col = cols[it.seg * d + j];
        unsigned long long m = 0;
        for (int64_t r = it.r0 + threadIdx.x; r < it.r1; r += 256) {
            const double x = b[r * ldb + col];
            if (isfinite(x)) {
                const unsigned long long bits = abs_bits(x);
                m = bits > m ? bits : m;
            }
        }
        m = block_max_u64(m, red);
        if (threadIdx.x == 0) atomicMax(&cmax[it.seg * d + j], m);
    }
}

// per (segment, column): the first node (caller's order) attaining the maximum
__global__ __launch_bounds__(256) void k_emb_first(const Item* __restrict__ items,
                                                   const double* __restrict__ b, int ldb,
                                                   const int32_t* __restrict__ rownode,
                                                   const int32_t* __restrict__ cols, int d,
                                                   const unsigned long long* __restrict__ cmax,
                                                   unsigned long long* __restrict__ first) {
    const Item it = items[blockIdx.x];
    for (int j = 0; j < d; ++j) {
        const int col = cols[it.seg * d + j];
        const unsigned long long target = cmax[it.seg * d + j];
        for (int64_t r = it.r0 + threadIdx.x; r < it.r1; r += 256) {
            const double x = b[r * ldb + col];
            if (isfinite(x) && abs_bits(x) == target) {
                const unsigned long long node = rownode ? (unsigned long long)rownode[r]
                                                        : (unsigned long long)r;
                atomicMin(&first[it.seg * d + j], node);
            }
        }
    }
}

// the entry at the first maximum (its sign decides the column's sign)
__global__ void k_emb_pick(int m, const double* __restrict__ b, int ldb,
                           const int32_t* __restrict__ rowmap, const int32_t* __restrict__ kind,
                           const int32_t* __restrict__ cols, int d,
                           const unsigned long long* __restrict__ first, double* __restrict__ val) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int c = t / d;
    if (kind[c] != 1 || first[t] == ~0ull) {
        val[t] = 0.0;
        return;
    }
    const int64_t node = (int64_t)first[t];
    const int64_t row = rowmap ? rowmap[node] : node;
    val[t] = b[row * ldb + cols[t]];
}

// coordinates in the caller's node order: kind 1 -> sign * x (then * f_c +
// meta_c when affine), kind 0 -> meta_c; and the bits of max |coordinate|
// (grid-stride, one atomic per workgroup)
#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void k_emb_coords(int64_t n, const double* __restrict__ b, int ldb,
                                                    const int32_t* __restrict__ rowmap,
                                                    const int32_t* __restrict__ lab,
                                                    const int32_t* __restrict__ kind,
                                                    const int32_t* __restrict__ cols, int d,
                                                    const double* __restrict__ sign,
                                                    const double* __restrict__ f,
                                                    const double* __restrict__ meta, int affine,
                                                    double* __restrict__ coords,
                                                    unsigned long long* __restrict__ gmax) {
    __shared__ unsigned long long red[4];
    const int64_t total = n * d;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int64_t v = t / d;
        const int j = (int)(t - v * d);
        const int c = lab ? lab[v] : 0;
        double y;
        if (kind[c] == 1) {
            const int64_t row = rowmap ? rowmap[v] : v;
            y = sign[c * d + j] * b[row * ldb + cols[c * d + j]];
            if (affine) {
                const double p = y * f[c];
                y = p + meta[c * d + j];
            }
        } else {
            y = meta[c * d + j];
        }
        coords[t] = y;
        if (isfinite(y)) {
            const unsigned long long bits = abs_bits(y);
            m = bits > m ? bits : m;
        }
    }
    m = block_max_u64(m, red);
    if (threadIdx.x == 0) atomicMax(gmax, m);
}

__global__ void k_emb_scale_noise(int64_t m, const double* __restrict__ coords, double ex,
                                  const double* __restrict__ noise, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const double scaled = coords[t] * ex;
    const double nz = 1e-4 * noise[t];
    out[t] = (float)scaled + (float)nz;
}
#pragma clang fp contract(on)

__global__ void k_emb_fill_u64(unsigned long long* __restrict__ a, int m, unsigned long long v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) a[t] = v;
}

inline dim3 grid_of(int64_t threads) { return dim3((unsigned)ceil_div(threads, 256)); }

}  // namespace

void embed_output_device(const EmbedPlan& p, hipStream_t s) {
    const int64_t n = p.n;
    const int d = p.d;
    const int nseg = p.nseg;
    const int m = nseg * d;
    SI_CHECK(n >= 1 && d >= 1 && nseg >= 1 && (nseg == 1 || p.lab), kErrInternal,
             "embed_output: bad plan");
    SI_CHECK((int)p.kind.size() == nseg && (int)p.cols.size() == m &&
                 (int)p.row0.size() == nseg && (int)p.rows.size() == nseg &&
                 (!p.affine || ((int)p.range.size() == nseg && (int)p.meta.size() == m)),
             kErrInternal, "embed_output: plan sizes");
    // work items: chunks of the row range of every segment with its own block
    std::vector<Item> items;
    for (int c = 0; c < nseg; ++c) {
        if (p.kind[c] != 1) continue;
        SI_CHECK(p.row0[c] >= 0 && p.rows[c] >= 1 && p.row0[c] + p.rows[c] <= n, kErrInternal,
                 "embed_output: segment rows");
        for (int64_t r = 0; r < p.rows[c]; r += kChunk)
            items.push_back(Item{p.row0[c] + r, p.row0[c] + std::min<int64_t>(p.rows[c], r + kChunk),
                                 c, 0});
    }
    // one device arena for the small arrays, laid out so that each host
    // transfer is one copy: [items | kind | cols] up, [cmax | val] down,
    // [sign | f | meta] up, gmax down
    auto al8 = [](size_t x) { return (x + 7) & ~(size_t)7; };
    const size_t ni = items.size();
    const size_t o_items = 0, o_kind = sizeof(Item) * ni;
    const size_t o_cols = al8(o_kind + 4 * (size_t)nseg), o_cmax = al8(o_cols + 4 * (size_t)m);
    const size_t o_val = o_cmax + 8 * (size_t)m, o_first = o_val + 8 * (size_t)m;
    const size_t o_sign = o_first + 8 * (size_t)m, o_f = o_sign + 8 * (size_t)m;
    const size_t o_meta = o_f + 8 * (size_t)nseg, o_gmax = o_meta + 8 * (size_t)m;
    const size_t bytes = o_gmax + 8;
    StreamBuf<char> arena(bytes, s);
    char* base = arena.get();
    const Item* ditems = reinterpret_cast<const Item*>(base + o_items);
    int32_t* kind = reinterpret_cast<int32_t*>(base + o_kind);
    int32_t* cols = reinterpret_cast<int32_t*>(base + o_cols);
    unsigned long long* cmax = reinterpret_cast<unsigned long long*>(base + o_cmax);
    double* val = reinterpret_cast<double*>(base + o_val);
    unsigned long long* first = reinterpret_cast<unsigned long long*>(base + o_first);
    double* sign = reinterpret_cast<double*>(base + o_sign);
    double* f = reinterpret_cast<double*>(base + o_f);
    double* meta = reinterpret_cast<double*>(base + o_meta);
    unsigned long long* gmax = reinterpret_cast<unsigned long long*>(base + o_gmax);
    std::vector<char> stage(bytes, 0);
    if (ni) std::memcpy(stage.data() + o_items, items.data(), sizeof(Item) * ni);
    std::memcpy(stage.data() + o_kind, p.kind.data(), 4 * (size_t)nseg);
    std::memcpy(stage.data() + o_cols, p.cols.data(), 4 * (size_t)m);
    to_device(base, stage.data(), o_cmax, s);
    hipLaunchKernelGGL(k_emb_fill_u64, grid_of(m), dim3(256), 0, s, cmax, m, 0ull);
    hipLaunchKernelGGL(k_emb_fill_u64, grid_of(m), dim3(256), 0, s, first, m, ~0ull);
    hipLaunchKernelGGL(k_emb_fill_u64, grid_of(1), dim3(256), 0, s, gmax, 1, 0ull);
    if (ni) {
        hipLaunchKernelGGL(k_emb_colmax, dim3((unsigned)ni), dim3(256), 0, s, ditems, p.b, p.ldb,
                           cols, d, cmax);
        hipLaunchKernelGGL(k_emb_first, dim3((unsigned)ni), dim3(256), 0, s, ditems, p.b, p.ldb,
                           p.rownode, cols, d, cmax, first);
    }
    hipLaunchKernelGGL(k_emb_pick, grid_of(m), dim3(256), 0, s, m, p.b, p.ldb, p.rowmap, kind,
                       cols, d, first, val);
    SI_LAUNCH_CHECK();
    to_host(stage.data() + o_cmax, base + o_cmax, o_first - o_cmax, s);
    const unsigned long long* hmax = reinterpret_cast<const unsigned long long*>(stage.data() + o_cmax);
    const double* hval = reinterpret_cast<const double*>(stage.data() + o_val);
    // per segment: column signs, and the factor range_c / max|sel|
    double* hs = reinterpret_cast<double*>(stage.data() + o_sign);
    double* hf = reinterpret_cast<double*>(stage.data() + o_f);
    for (int t = 0; t < m; ++t) hs[t] = 1.0;
    for (int c = 0; c < nseg; ++c) hf[c] = 1.0;
    for (int c = 0; c < nseg; ++c) {
        if (p.kind[c] != 1) continue;
        double mx = 0.0;
        for (int j = 0; j < d; ++j) {
            hs[c * d + j] = std::signbit(hval[c * d + j]) ? -1.0 : 1.0;
            double a;
            std::memcpy(&a, &hmax[c * d + j], sizeof(a));
            mx = std::max(mx, a);
        }
        if (p.affine) hf[c] = mx > 0.0 ? p.range[c] / mx : 1.0;
    }
    if (p.affine) std::memcpy(stage.data() + o_meta, p.meta.data(), 8 * (size_t)m);
    to_device(base + o_sign, stage.data() + o_sign, o_gmax - o_sign, s);
    double* coords = ws_get<double>(kWsAbiIn, (size_t)n * d);
    const unsigned cblocks = (unsigned)std::min<int64_t>(kMaxBlocks, ceil_div(n * d, 256));
    hipLaunchKernelGGL(k_emb_coords, dim3(cblocks), dim3(256), 0, s, n, p.b, p.ldb, p.rowmap, p.lab,
                       kind, cols, d, sign, f, meta, p.affine ? 1 : 0, coords, gmax);
    SI_LAUNCH_CHECK();
    unsigned long long gbits = 0;
    to_host(&gbits, gmax, 1, s);
    double mx;
    std::memcpy(&mx, &gbits, sizeof(mx));
    if (!(mx > 0.0))
        throw Error(kErrInvalidGraph,
                    "invalid graph: scale_coords: embedding is all-zero; cannot scale to max=10");
    const double ex = 10.0 / mx;
    double* nz = ws_get<double>(kWsAbiOut, (size_t)n * d);
    fill_normal_f64(nz, n, d, d, p.seed, kStreamNoise, s);
    float* out = ws_get<float>(kWsResR, (size_t)n * d);
    hipLaunchKernelGGL(k_emb_scale_noise, grid_of(n * d), dim3(256), 0, s, n * d, coords, ex, nz,
                       out);
    SI_LAUNCH_CHECK();
    download_to_host(p.out, out, sizeof(float) * (size_t)n * d, s);
}

}  // namespace si
