// Shared definitions for the VariantDBSCAN MI355X library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace vdb {

constexpr int kWave = 64;

// ------------------------------------------------------------------ errors
// Every library entry point returns an int status; 0 = success.
enum Status : int {
    kOk = 0,
    kErrInvalid = 1,     // bad argument (NULL pointer, MBBsize <= 0, ...)
    kErrNonFinite = 2,   // NaN/inf coordinate
    kErrDevice = 3,      // HIP runtime failure
    kErrNoMemory = 4,    // device or host allocation failed
    kErrUnsupported = 5, // request outside what this build handles
};

void set_error(const std::string &msg);
const char *last_error();

#define VDB_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::vdb::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + \
                             " @" + __FILE__ + ":" + std::to_string(__LINE__)); \
            return e_ == hipErrorOutOfMemory ? ::vdb::kErrNoMemory             \
                                             : ::vdb::kErrDevice;              \
        }                                                                      \
    } while (0)

#define VDB_TRY(expr)                                                          \
    do {                                                                       \
        int s_ = (expr);                                                       \
        if (s_ != ::vdb::kOk) return s_;                                       \
    } while (0)

// --------------------------------------------------------- exact predicate
// The reference accepts q as an eps-neighbour of p when
//   sqrt((px-qx)*(px-qx) + (py-qy)*(py-qy)) <= eps        (src/DBScan.cpp:687,699-702)
// evaluated in binary64 with every operation rounded separately. sqrt_rn is
// monotone, so this equals  s <= T(eps)  with s the unrounded-free sum below
// and T(eps) the largest double whose rounded square root is <= eps
// (computed on the host, threshold_for_eps). The explicit _rn intrinsics keep
// the compiler from contracting the products into an FMA.
__device__ __forceinline__ double dist2_exact(double px, double py, double qx,
                                              double qy)
{
    double dx = __dsub_rn(px, qx);
    double dy = __dsub_rn(py, qy);
    return __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
}

// Exact lower / upper bounds of dist2_exact(p, q) over every q inside a box
// [bx0,bx1]x[by0,by1] (bounds taken from actual point coordinates). Rounding
// is monotone in each operand, so evaluating the formula at the extreme
// coordinate differences bounds every pair inside the box.
__device__ __forceinline__ double box_dist2_min(double px, double py, double bx0,
                                                double bx1, double by0, double by1)
{
    double gx = 0.0, gy = 0.0;
    if (px < bx0) gx = __dsub_rn(bx0, px);
    else if (px > bx1) gx = __dsub_rn(px, bx1);
    if (py < by0) gy = __dsub_rn(by0, py);
    else if (py > by1) gy = __dsub_rn(py, by1);
    return __dadd_rn(__dmul_rn(gx, gx), __dmul_rn(gy, gy));
}

__device__ __forceinline__ double box_dist2_max(double px, double py, double bx0,
                                                double bx1, double by0, double by1)
{
    double gx = fmax(__dsub_rn(px, bx0), __dsub_rn(bx1, px));
    double gy = fmax(__dsub_rn(py, by0), __dsub_rn(by1, py));
    return __dadd_rn(__dmul_rn(gx, gx), __dmul_rn(gy, gy));
}

// Exact lower / upper bounds of dist2_exact(p, q) over every p in box a and q
// in box b (boxes {x0, x1, y0, y1} of actual point coordinates), by the same
// monotonicity argument.
__device__ __forceinline__ double boxes_dist2_min(const double *a, const double *b)
{
    double gx = 0.0, gy = 0.0;
    if (b[0] > a[1]) gx = __dsub_rn(b[0], a[1]);
    else if (a[0] > b[1]) gx = __dsub_rn(a[0], b[1]);
    if (b[2] > a[3]) gy = __dsub_rn(b[2], a[3]);
    else if (a[2] > b[3]) gy = __dsub_rn(a[2], b[3]);
    return __dadd_rn(__dmul_rn(gx, gx), __dmul_rn(gy, gy));
}

__device__ __forceinline__ double boxes_dist2_max(const double *a, const double *b)
{
    double gx = fmax(__dsub_rn(b[1], a[0]), __dsub_rn(a[1], b[0]));
    double gy = fmax(__dsub_rn(b[3], a[2]), __dsub_rn(a[3], b[2]));
    return __dadd_rn(__dmul_rn(gx, gx), __dmul_rn(gy, gy));
}

// --------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// number of set bits of m strictly below this lane
__device__ __forceinline__ int mask_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ int wave_id_in_block() { return threadIdx.x / kWave; }

template <typename T> __device__ __forceinline__ T uniform(T v)
{
    // value of lane 0 (caller guarantees the value is wave-uniform)
    return __shfl(v, 0);
}

// v of lane (lane ^ J), J a constant power of two below 64, without the LDS
// permute: DPP moves inside rows of 16 lanes (J = 1, 2: quad permutes; 4:
// quad reversal then row half-mirror; 8: row rotate by 8) and the gfx950
// row swaps across rows (J = 16, 32: permlane16/32_swap of v with itself,
// the lane taking the half that came from the other row).
template <int J> __device__ __forceinline__ uint32_t lane_xor(uint32_t v)
{
    static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "xor mask");
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false);
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false);
    } else if constexpr (J == 4) {
        int t = __builtin_amdgcn_update_dpp(0, (int)v, 0x1B, 0xf, 0xf, false);
        return (uint32_t)__builtin_amdgcn_update_dpp(0, t, 0x141, 0xf, 0xf, false);
    } else if constexpr (J == 8) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false);
    } else if constexpr (J == 16) {
        auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (__lane_id() & 16) ? r[0] : r[1];
    } else {
        auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (__lane_id() & 32) ? r[0] : r[1];
    }
}
// v of lane - 1 (lane 0 keeps its own): the DPP whole-wave shift
__device__ __forceinline__ uint32_t lane_up1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ double lane_up1_f64(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    uint32_t lo = lane_up1((uint32_t)b), hi = lane_up1((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int J> __device__ __forceinline__ double lane_xor_f64(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    uint32_t lo = lane_xor<J>((uint32_t)b), hi = lane_xor<J>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// lanes of the wave holding the same key d < 2^BITS (d == 2^BITS: no key, a
// lane that matches no key lane), from BITS + 1 ballots
template <int BITS> __device__ __forceinline__ uint64_t bucket_peers(uint32_t d)
{
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b <= BITS; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bal = ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

} // namespace vdb
