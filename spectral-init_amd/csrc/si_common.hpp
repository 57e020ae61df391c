// si_common.hpp -- shared host/device definitions of the MI355X spectral
// initialisation library (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace si {

// Error carried from the HIP runtime or from an internal check to the C ABI,
// which converts it to a status code (exceptions never cross the ABI).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : int {
    kOk = 0,
    kErrConvergence = 1,
    kErrInvalidGraph = 2,
    kErrTooFewNodes = 3,
    kErrHip = 4,
    kErrInternal = 5,
    kErrInvalidArgument = 6,
    kErrUnsupported = 7,
};

#define SI_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::si::Error(::si::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                 " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define SI_CHECK(cond, code, msg)                       \
    do {                                                \
        if (!(cond)) throw ::si::Error((code), (msg));  \
    } while (0)

#define SI_LAUNCH_CHECK() SI_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Philox4x32-10 + Box-Muller normal; identical arithmetic to the oracle's
// counter RNG so that device and CPU solves can start from the same block.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__host__ __device__ inline double normal_at(uint64_t seed, uint32_t stream, uint64_t index) {
    uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x5eedu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double inv53 = 1.0 / 9007199254740992.0;
    uint64_t a = (((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11)) & ((1ull << 53) - 1);
    uint64_t b = (((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11)) & ((1ull << 53) - 1);
    double u1 = ((double)a + 1.0) * inv53;
    double u2 = (double)b * inv53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// The FP32 start blocks' stream: one Philox4x32-10 call per (row, column
// pair) gives two normals, the cosine and sine outputs of one Box-Muller
// transform in FP32 (24-bit uniforms): half the Philox calls of normal_at and
// no FP64 transcendentals.  (counter = pair index; z0 -> column 2p, z1 -> 2p+1.)
__host__ __device__ inline void normal_pair_f32(uint64_t seed, uint32_t stream, uint64_t index,
                                                float& z0, float& z1) {
    uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x9a1fu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float inv24 = 1.0f / 16777216.0f;
    const float u1 = ((float)(c[0] >> 8) + 1.0f) * inv24;  // (0, 1]
    const float u2 = (float)(c[1] >> 8) * inv24;           // [0, 1)
    const float r = sqrtf(-2.0f * logf(u1));
    const float t = 6.2831853071795865f * u2;
    z0 = r * cosf(t);
    z1 = r * sinf(t);
}

constexpr uint32_t kStreamLobpcgInit = 1u;
constexpr uint32_t kStreamRsvdOmega = 2u;
constexpr uint32_t kStreamNoise = 7u;
constexpr uint32_t kStreamProbe = 9u;     // warm-start network: random-walk probes
constexpr uint32_t kStreamNetInit = 11u;  // warm-start network: initial weights
constexpr uint32_t kStreamDenseInvit = 13u;  // dense eigensolver: inverse-iteration starts
constexpr uint32_t kStreamOrder = 15u;       // clustering locality order: walk start

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// SI_TRACE=1 prints diagnostics to stderr (solver rounds, build phases).
inline bool trace_on() {
    static const bool on = [] {
        const char* v = std::getenv("SI_TRACE");
        return v && v[0] == '1';
    }();
    return on;
}

}  // namespace si
