// k_batch.hip -- the round planner of the batched component solver, on the
// device.
//
// The batched solver (solver_batch.cpp) runs rounds of "Chebyshev filter of
// per-segment degree, then segmented Rayleigh-Ritz" over every connected
// component of a block-diagonal Laplacian.  The decisions between rounds --
// which segments have converged, stalled or failed, each segment's damped
// interval [a_c, 2] and filter degree, the per-step recurrence scalars of every
// segment -- are made here, by one workgroup that reads the round's Ritz
// values, residual norms and Cholesky status where the Rayleigh-Ritz kernels
// left them.  It writes:
//   * the coefficient table of the next round's filter steps,
//   * the gate ctl[0]: the number of filter steps the next round runs (0 = no
//     round: every segment is done or the round limit is reached).
// Every launch of a round reads the gate when it starts: filter step j runs
// only while j <= gate, the Rayleigh-Ritz launches only while gate > 0.  So a
// captured graph of several rounds replays without the host in between, and
// rounds after the last useful one cost only their empty launches.
//
// The filter steps take their buffers from a three-buffer ring by the gate
// (spmm_seg_ring_*), so the last step that runs always writes the buffer the
// captured Rayleigh-Ritz reads.
#include "si_device.hpp"

namespace si {

namespace {

constexpr int kPlanThreads = 256;

// Largest degree d <= cap with T_d(c/e) <= e^9.2 on [lo, b] (the
// amplification cap of the single-graph solver).
__device__ int plan_degree(double lo, double b, int cap) {
    const double c = (lo + b) / 2.0, e = (b - lo) / 2.0;
    const double x0 = c / e;
    const int dmax = (int)floor(9.2 / acosh(x0 > 1.0 + 1e-12 ? x0 : 1.0 + 1e-12));
    const int d = cap < dmax ? cap : dmax;
    return d > 2 ? d : 2;
}

// max_j ||L v_j - theta_j v_j|| / max(||v_j||, eps) of segment c from the
// squared norms of the Rayleigh-Ritz update; a NaN anywhere gives +inf.
__device__ double plan_worst(const double* __restrict__ norms2, int c, int k) {
    double wv = 0.0;
    for (int j = 0; j < k; ++j) {
        const double r2 = norms2[(int64_t)c * 2 * k + j];
        const double x2 = norms2[(int64_t)c * 2 * k + k + j];
        if (r2 != r2 || x2 != x2) return INFINITY;
        double xn = sqrt(x2 > 0.0 ? x2 : 0.0);
        if (xn < 2.220446049250313e-16) xn = 2.220446049250313e-16;
        const double r = sqrt(r2 > 0.0 ? r2 : 0.0) / xn;
        wv = r > wv ? r : wv;
    }
    return wv;
}

// mode 0: FP32 rounds (a segment stops at tol, at a 1e-4 stall floor, or on a
//         failed Cholesky; a_c from its latest Ritz values);
// mode 1: FP64 rounds (a segment stops at tol or on a failed Cholesky; its
//         kept Ritz values thk and a_c follow every round it is re-projected).
template <typename T>
__global__ __launch_bounds__(kPlanThreads) void k_batch_plan(BatchPlan p, int mode, int update,
                                                             T* __restrict__ coef) {
    __shared__ int s_dmax, s_active;
    const int t = threadIdx.x;
    const int prev_gate = p.ctl[0];  // read by every thread before thread 0 rewrites it
    if (t == 0) { s_dmax = 0; s_active = 0; }
    __syncthreads();
    // update 1: the round just planned by the gate ran if the gate was open;
    // update 2: results of a round run without a gate (always read)
    const bool ran = update == 2 || (update == 1 && prev_gate > 0);
    const int k = p.k, m = p.m;
    for (int c = t; c < p.nseg; c += kPlanThreads) {
        int done = p.done[c];
        if (ran) {
            if (mode == 0) {
                if (done == 0) {
                    if (p.status[c] != 0) {
                        done = 2;  // Y^T Y not positive definite
                    } else {
                        const double wv = plan_worst(p.norms2, c, k);
                        p.worst[c] = wv;
                        if (wv <= p.tol) {
                            done = 1;
                        } else if (wv < 0.7 * p.best[c]) {
                            p.best[c] = wv;
                            p.stall[c] = 0;
                        } else if (++p.stall[c] >= 3 && wv < 1e-4) {
                            done = 1;  // FP32 floor
                        }
                        if (done == 0) {
                            const double* tc = p.theta + (int64_t)c * m;
                            const double lo = tc[k - 1] * 1.5 + 1e-12;
                            const double hi = tc[m - 1] > lo ? tc[m - 1] : lo;
                            p.a[c] = hi < 0.6 ? hi : 0.6;
                        }
                    }
                }
            } else if (done != 2) {
                if (p.status[c] != 0) {
                    done = 2;
                    p.worst[c] = INFINITY;
                } else {
                    for (int j = 0; j < m; ++j)
                        p.thk[(int64_t)c * m + j] = p.theta[(int64_t)c * m + j];
                    const double wv = plan_worst(p.norms2, c, k);
                    p.worst[c] = wv;
                    if (wv <= p.tol) done = 1;
                }
            }
            p.done[c] = done;
        }
        if (mode == 1 && done == 0) {
            const double* tc = p.thk + (int64_t)c * m;
            const double lo = tc[k - 1] * 1.5 + 1e-12;
            const double hi = tc[m - 1] > lo ? tc[m - 1] : lo;
            p.a[c] = hi < 0.6 ? hi : 0.6;
        }
        const int deg = done ? 0 : plan_degree(p.a[c], p.b, p.cap < p.steps ? p.cap : p.steps);
        p.deg[c] = deg;
        if (deg > 0) {
            atomicMax(&s_dmax, deg);
            atomicAdd(&s_active, 1);
        }
    }
    __syncthreads();
    int gate = s_dmax;
    if (p.ctl[1] >= p.max_rounds) gate = 0;
    // the recurrence scalars of steps 1..steps (row nseg: rows of no segment)
    const int rows = p.nseg + 1;
    for (int e = t; e < p.steps * rows; e += kPlanThreads) {
        const int j = e / rows + 1, c = e % rows;
        T* q = coef + (int64_t)e * 3;
        const int deg = c < p.nseg ? p.deg[c] : 0;
        if (j <= deg) {
            const double lo = p.a[c];
            const double cc = (lo + p.b) / 2.0, ee = (p.b - lo) / 2.0;
            if (j == 1) { q[0] = (T)(1.0 / ee); q[1] = (T)(-cc / ee); q[2] = T(0); }
            else { q[0] = (T)(2.0 / ee); q[1] = (T)(-2.0 * cc / ee); q[2] = T(-1); }
        } else {
            q[0] = T(0); q[1] = T(1); q[2] = T(0);  // copy through
        }
    }
    __syncthreads();
    if (t == 0) {
        p.ctl[0] = gate;
        if (gate > 0) {
            // filter steps + the Rayleigh-Ritz product; an FP32 round after the
            // first starts from the known L X (no SpMM for its first step)
            p.ctl[2] += gate + 1 - ((mode == 0 && p.ctl[1] > 0) ? 1 : 0);
            p.ctl[1] += 1;
        }
        p.ctl[3] = s_active;
    }
}

__global__ void k_batch_init(BatchPlan p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < p.nseg) {
        p.a[c] = 0.1;
        p.best[c] = INFINITY;
        p.worst[c] = INFINITY;
        p.stall[c] = 0;
        p.done[c] = 0;
        p.deg[c] = 0;
        if (p.thk)
            for (int j = 0; j < p.m; ++j) p.thk[(int64_t)c * p.m + j] = 0.0;
    }
    if (c < 4) p.ctl[c] = 0;
}

// thread per row: the last segment starting at or before the row (binary
// search over the sorted starts) holds it if the row is before its end
__global__ void k_segment_map(int32_t* __restrict__ rseg, int64_t n, int nseg,
                              const int64_t* __restrict__ starts, const int64_t* __restrict__ ends,
                              const int32_t* __restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nseg;  // first segment with starts > i is at hi
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (starts[mid] <= i) lo = mid + 1;
        else hi = mid;
    }
    const int j = lo - 1;
    rseg[i] = (j >= 0 && i < ends[j]) ? ids[j] : nseg;
}

// Row range of every component in the operator's row order: rows [rr[c],
// rr[nc + c]] carry component c (perm[row] = input node, comp = its label).
// Only the ends of runs write (a component in one contiguous run writes each
// end once); rr starts at (INT32_MAX, -1).
__global__ void k_component_rows(int64_t n, const int32_t* __restrict__ perm,
                                 const int32_t* __restrict__ comp, int32_t nc,
                                 int32_t* __restrict__ rr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t c = comp[perm[r]];
    if (r == 0 || comp[perm[r - 1]] != c) atomicMin(&rr[c], (int32_t)r);
    if (r == n - 1 || comp[perm[r + 1]] != c) atomicMax(&rr[nc + c], (int32_t)r);
}

__global__ void k_component_rows_init(int32_t nc, int32_t* __restrict__ rr) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < nc) rr[i] = INT32_MAX;
    else if (i < 2 * nc) rr[i] = -1;
}

// First filter step of a round from the known L X (ax, from the previous
// Rayleigh-Ritz update): out = s_l * ax + s_x * x per row with the segment's
// step-1 scalars (s_l = 0: a copy-through segment, ax not read), written to the
// ring buffer that spmm_seg_ring's step 1 would write (ring[(g - 1) mod 3]).
// Thread per 4 columns; returns at once when the gate is 0.
__global__ void k_axpby_seg_ring(int64_t n, int w4, int ld, const int32_t* __restrict__ rseg,
                                 const float* __restrict__ coef, const float* __restrict__ ax,
                                 const float* __restrict__ x, float* r0, float* r1, float* r2,
                                 const int32_t* __restrict__ gate) {
    const int g = *gate;
    if (g < 1) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * w4) return;
    const int i0 = ((g - 1) % 3 + 3) % 3;
    float* out = i0 == 0 ? r0 : (i0 == 1 ? r1 : r2);
    const int64_t i = e / w4;
    const int64_t o = i * ld + 4 * (e % w4);
    const int c = rseg[i];
    const float sl = coef[3 * c], sx = coef[3 * c + 1];
    const float4 xv = *reinterpret_cast<const float4*>(x + o);
    float4 r;
    if (sl == 0.f) {
        r = make_float4(sx * xv.x, sx * xv.y, sx * xv.z, sx * xv.w);
    } else {
        const float4 av = *reinterpret_cast<const float4*>(ax + o);
        r = make_float4(sl * av.x + sx * xv.x, sl * av.y + sx * xv.y, sl * av.z + sx * xv.z,
                        sl * av.w + sx * xv.w);
    }
    *reinterpret_cast<float4*>(out + o) = r;
}

// workgroup per work item: its rows [blk_rows[2b], blk_rows[2b+1]), k columns
__global__ void k_copy_item_rows(const int64_t* __restrict__ blk_rows,
                                 const double* __restrict__ src, int lds, double* __restrict__ dst,
                                 int ldd, int k) {
    const int64_t r0 = blk_rows[2 * blockIdx.x], r1 = blk_rows[2 * blockIdx.x + 1];
    const int64_t cnt = (r1 - r0) * k;
    for (int64_t e = threadIdx.x; e < cnt; e += blockDim.x) {
        const int64_t r = r0 + e / k;
        const int c = (int)(e % k);
        dst[r * ldd + c] = src[r * lds + c];
    }
}

}  // namespace

void batch_axpby_ring_f32(int64_t n, int w, int ld, const int32_t* rseg, const float* coef,
                          const float* ax, const float* x, float* const ring[3],
                          const int32_t* gate, hipStream_t s) {
    SI_CHECK(w % 4 == 0 && ld % 4 == 0 && ld >= w, kErrInvalidArgument,
             "batch axpby: widths must be multiples of 4");
    SI_CHECK((reinterpret_cast<uintptr_t>(ax) | reinterpret_cast<uintptr_t>(x) |
              reinterpret_cast<uintptr_t>(ring[0]) | reinterpret_cast<uintptr_t>(ring[1]) |
              reinterpret_cast<uintptr_t>(ring[2])) % 16 == 0,
             kErrInvalidArgument, "batch axpby: 16-byte aligned blocks required");
    const int64_t threads = n * (w / 4);
    if (threads <= 0) return;
    hipLaunchKernelGGL(k_axpby_seg_ring, dim3((unsigned)ceil_div(threads, (int64_t)256)), dim3(256),
                       0, s, n, w / 4, ld, rseg, coef, ax, x, ring[0], ring[1], ring[2], gate);
    SI_LAUNCH_CHECK();
}

void batch_copy_rows(const SegWork& items, const double* src, int lds, double* dst, int ldd,
                     int k, hipStream_t s) {
    if (items.nblk <= 0) return;
    hipLaunchKernelGGL(k_copy_item_rows, dim3((unsigned)items.nblk), dim3(256), 0, s,
                       items.blk_rows, src, lds, dst, ldd, k);
    SI_LAUNCH_CHECK();
}

void batch_segment_map(int32_t* rseg, int64_t n, int nseg, const int64_t* starts,
                       const int64_t* ends, const int32_t* ids, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_segment_map, dim3((unsigned)ceil_div(n, (int64_t)256)), dim3(256), 0, s,
                       rseg, n, nseg, starts, ends, ids);
    SI_LAUNCH_CHECK();
}

void component_rows(int64_t n, const int32_t* perm, const int32_t* comp, int64_t nc, int32_t* rr,
                    hipStream_t s) {
    SI_CHECK(n >= 1 && nc >= 1 && nc <= n && n < (int64_t)INT32_MAX, kErrInvalidArgument,
             "component rows: bad sizes");
    hipLaunchKernelGGL(k_component_rows_init, dim3((unsigned)ceil_div(2 * nc, (int64_t)256)),
                       dim3(256), 0, s, (int32_t)nc, rr);
    hipLaunchKernelGGL(k_component_rows, dim3((unsigned)ceil_div(n, (int64_t)256)), dim3(256), 0, s,
                       n, perm, comp, (int32_t)nc, rr);
    SI_LAUNCH_CHECK();
}

void batch_plan_init(const BatchPlan& p, hipStream_t s) {
    SI_CHECK(p.nseg >= 1, kErrInvalidArgument, "batch plan: no segments");
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)ceil_div((int64_t)p.nseg + 4, (int64_t)256)),
                       dim3(256), 0, s, p);
    SI_LAUNCH_CHECK();
}

void batch_plan_f32(const BatchPlan& p, int update, float* coef, hipStream_t s) {
    SI_CHECK(p.steps >= 2 && p.k >= 1 && p.k <= p.m, kErrInvalidArgument, "batch plan: bad sizes");
    hipLaunchKernelGGL((k_batch_plan<float>), dim3(1), dim3(kPlanThreads), 0, s, p, 0, update, coef);
    SI_LAUNCH_CHECK();
}

void batch_plan_f64(const BatchPlan& p, int update, double* coef, hipStream_t s) {
    SI_CHECK(p.steps >= 2 && p.k >= 1 && p.k <= p.m && p.thk, kErrInvalidArgument,
             "batch plan: bad sizes");
    hipLaunchKernelGGL((k_batch_plan<double>), dim3(1), dim3(kPlanThreads), 0, s, p, 1, update,
                       coef);
    SI_LAUNCH_CHECK();
}

}  // namespace si
