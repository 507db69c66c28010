// Device-side shared structures and helpers for the floodhip kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>
#include <stdexcept>
#include <string>

namespace fh {

// Argument errors raised by host-side setup (mapped to FH_ERR_VALUE).
struct ValueError {
    std::string msg;
};

template <typename real> struct V4;
template <> struct V4<float> { typedef float4 type; };
template <> struct V4<double> { typedef double4 type; };

constexpr int kLedgerTerms = 13;
// MassLedger order (parity_runner.py:38-51)
enum LedgerTerm {
    L_RAIN = 0, L_INFIL, L_BNET, L_BIN, L_BOUT, L_EXTERNAL, L_CAPTURE, L_SURCHARGE,
    L_OUTFALL, L_OVERFLOW, L_UNRETURNED, L_RCAPTURE, L_RLOSS
};

// Start-of-step control state of one member, double-buffered by step parity:
// every block of a step's update kernel reads clk[p] and the update's first
// block writes clk[1 - p] for the next step, so no block ever reads a clock
// that another block of the same launch is writing.
// The rain interval of a time t: the searchsorted(end, t, side="right")
// position `hint` of parity_runner.py:96-102, the start and rate of that
// interval (start +inf when past the last one) and its end, the time up to which
// the same position holds (sorted end times; -inf: look the position up).
struct RainCache {
    double next_end, start, rate;
    int32_t hint, pad;
};

struct Clock {
    double t;              // FP64 simulated time at the start of the step
    double cached_cfl;     // last evaluated CFL dt (state precision value)
    double dt_prev;        // dt of the previous executed step (its pending wet-time increment)
    RainCache rain;        // rain interval found for t
    int64_t step;          // global step counter (steps executed)
    int64_t interval_step; // steps executed in the current advance call
    int64_t cfl_evals;
    int32_t eval;          // this step evaluates the CFL condition
    int32_t active;        // this step executes: t < target && interval_step < max_steps
    int32_t pending_wet;   // the previous executed step's wet-time increment is pending
    // inflow-table search hint: the table position found at an earlier time of
    // this clock (reset whenever the clock or the forcing is set); used only when
    // the table's times are sorted, which the host checks
    int32_t ext_hint;
};

// Step control block of one member, resident in device memory.  The clock of
// the next step is written only by the first block of the update kernel; the
// outputs of the current step (dt, rain rate, inflow interpolation, active) are
// written there too, for the drainage kernels and the host; ledger terms are
// written only by k_fold.
struct Ctl {
    Clock clk[2];
    int64_t li_cfl_step[2];           // step whose local-inertial CFL pass filled that slot (-1: none)
    double target;         // stop when t >= target
    double final_t;        // event end time (dt cap final_t - t)
    int64_t max_steps;     // step budget of the current advance call
    double override_dt_value;
    double override_rain;
    int32_t override_dt;   // host-driven single step: dt and rain given by the host
    int32_t rain_sorted, ext_sorted;
    // outputs of the current step
    int32_t active;        // the current step executes
    double dt;             // dt of the current step, as applied (== (double)dt_real)
    double dt_real;        // dt of the current step in state precision (stored as double)
    double rain_rate;      // rain rate of the current step (state precision value)
    double t_step;         // start time of the current step
    double ext_frac;       // external inflow interpolation fraction at t + dt/2
    int32_t ext_lower, ext_upper;
    int32_t nonfinite;     // a non-finite depth or node volume was found (k_fold's check)
    int32_t li_missed;     // a local-inertial step evaluated the CFL condition without its CFL pass
    double cfl_min;        // result of the last CFL reduction
    double diag[kLedgerTerms];    // this step's ledger increments
    double ledger[kLedgerTerms];  // cumulative FP64 ledger
};

// Block -> member mapping of one index space (cells, edges, nodes, ...).  Each
// member's items start on a tile boundary (tile = blockDim.x * items per
// thread), so a block never straddles two members.
struct Seg {
    const int* blk_member;   // [blocks]   member of each (logical) block
    const int* mem_blk0;     // [members]  first block of the member
    const int* mem_nblk;     // [members]  number of blocks of the member
    const int* mem_end;      // [members]  exclusive end of the member's real items
    // Shared-mesh batch (sh_members > 0): every member has the same mesh, whose
    // tables are stored once with member-local indices.  Member m's items are
    // [m * sh_items, (m + 1) * sh_items) (sh_nblk blocks each), and sh_link is the
    // per-member stride of the space this one indexes (cells for edges, edges
    // for cells).  sh_recip = floor(2^40 / sh_nblk) + 1 divides a block index by
    // sh_nblk exactly (the host checks blocks * sh_nblk < 2^40); sh_rot staggers
    // the members' tile orders.
    int sh_members = 0, sh_nblk = 0, sh_items = 0, sh_link = 0, sh_rot = 0;
    unsigned long long sh_recip = 0;
};

// Block placement of a shared-mesh batch.  Members run one after another (slot
// i belongs to member m = i / sh_nblk), and member m starts its tiles at
// m * sh_rot (mod sh_nblk), so the XCDs, each working a contiguous range of
// slots, read different parts of the shared tables at any moment instead of
// all requesting the same lines.  Returns the block b = m * sh_nblk + t, whose
// items, partial slots and member are those of the concatenated layout.  All
// of it is scalar arithmetic on the block index.
__device__ __forceinline__ int shared_block(const Seg& s, int i, int& m) {
    m = (int)(((unsigned long long)(unsigned)i * s.sh_recip) >> 40);
    int t = i - m * s.sh_nblk + m * s.sh_rot;
    if (t >= s.sh_nblk) t -= s.sh_nblk;
    return m * s.sh_nblk + t;
}

// CFL minimum of a step, spread over kCflSlots words of one 128-byte line each,
// [member][parity][slot]: every block adds its block minimum with one atomic min
// into slot (block index mod kCflSlots), so about 1/kCflSlots of the blocks meet
// on each address instead of all of them on one (device-scope atomics on one
// address serialize); readers take the minimum of the slots.  Idle slots hold
// +inf: a step's committer resets the next step's slots.
constexpr int kCflSlots = 64;
constexpr int kCflStride = 16;     // words per slot: one 128-byte line
__device__ __forceinline__ unsigned long long* cfl_slot(unsigned long long* base, int m, int parity, int k) {
    return base + ((size_t)(m * 2 + parity) * kCflSlots + k) * kCflStride;
}

template <typename real>
struct KP {                 // kernel constants, pre-rounded to the state precision
    real g, eps, half_g, two_g, cfl, max_speed, tiny, capture_q, pow43, reuse, slot_coef;
    double pow43_dln2;       // ((double)pow43 - 4/3) ln 2, for fric_pow in float
    double max_dt, dt_floor, fixed_dt;
    double g_d;              // gravity in double (products of two host scalars, g * dt)
    int interval;
    int remap;
    unsigned long long* cfl_slots;   // the CFL minimum slots (cfl_slot)
};

// A local-inertial edge record {Q, Q nx, Q ny}: three values, no padding.
template <typename real>
struct Rec3 {
    real x, y, z;
};

// NaN-propagating max/min, like numpy.maximum / jax.numpy.maximum (fmax/fmin
// would drop a NaN operand and hide a blown-up state from the finite checks).
// In float they are gfx950's v_maximum3_f32 / v_minimum3_f32 (IEEE 754-2019
// maximum / minimum), one instruction against a compare pair and a select; the
// one difference from the select form is the sign of a zero result when the
// operands are +0 and -0 (those order -0 below +0).
template <typename T>
__device__ __forceinline__ T vmax(T a, T b) { return (a != a || a > b) ? a : b; }
template <typename T>
__device__ __forceinline__ T vmin(T a, T b) { return (a != a || a < b) ? a : b; }
template <>
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_elementwise_maximum(a, b); }
template <>
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_elementwise_minimum(a, b); }

// Division by a positive divisor b, rounded to nearest-even exactly as IEEE
// a / b, with the reciprocal computed once per divisor and shared by every
// dividend.  In float: y0 = v_rcp_f32(b) (within 1 ulp), one Newton step
// y = y0 + y0 (1 - b y0) gives y = RN(1/b) (checked on the device for every
// significand, fh_division_selftest), then q = RN(a y), the remainder
// a - b q is exact by FMA, and RN(q + (a - b q) y) = RN(a / b) (Markstein's
// theorem: y the correctly rounded reciprocal, q within an ulp of a / b).  The
// remainder is formed as -(b q - a), so a zero dividend keeps its sign as in
// IEEE.  Exact for 2^-100 <= b <= 2^100 and every finite a with |a| >= 2^-100 or
// a = 0, which covers the kernels' divisors (floored at the wet threshold or
// 1e-12, or >= 1); below 2^-100 the remainder can underflow and the quotient
// may be one ulp off.  About 4 instructions per dividend and 3 per divisor,
// against 10 for each IEEE division sequence (two scales, reciprocal, five
// FMAs, fmas and fixup).  In double it is the plain division.
template <typename real>
struct Divisor {
    real b;
    __device__ __forceinline__ explicit Divisor(real b_) : b(b_) {}
    __device__ __forceinline__ real div(real a) const { return a / b; }
};
template <>
struct Divisor<float> {
    float b, y;
    __device__ __forceinline__ explicit Divisor(float b_) : b(b_) {
        const float y0 = __builtin_amdgcn_rcpf(b);
        y = __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
    }
    __device__ __forceinline__ float div(float a) const {
        const float q = a * y;
        const float t = __builtin_fmaf(b, q, -a);   // -(a - b q), exact
        return __builtin_fmaf(-t, y, q);
    }
};

// x^p for the friction exponent p = (float)(4/3) = 4/3 + d, d = 3.97e-8, as a
// float32 program raises to "4/3", for x >= 2^-100 (the callers floor their
// depth or radius).  x^p = (x r)^2 (1 + d ln x) with r = x^(-1/3): r0 from the
// hardware log2 and exp2 (relative error ~1e-6), one Newton step
// r = r0 (4 - x r0^3) / 3 in FP64 (error ~2 e0^2), then the product in FP64
// (x^d - 1 - d ln x is below 1e-13) and one rounding to float.  That is the
// correctly rounded x^p unless x^p lies within ~1e-12 relative of a rounding
// boundary: over every float in [2^-40, 2^20) (fh_pow_selftest) it is one ulp
// off for 4,284 of 503 M values, the device library's powf for 127 M.  About 17
// instructions against 137.  FP64 runs keep pow.
__device__ __forceinline__ float fric_pow(float x, double dln2) {
    const float l2 = __builtin_amdgcn_logf(x);
    const double r0 = (double)__builtin_amdgcn_exp2f(l2 * (-1.0f / 3.0f));
    const double xd = (double)x;
    const double r = r0 * (4.0 - xd * (r0 * r0 * r0)) * (1.0 / 3.0);
    const double t = xd * r;
    return (float)((t * t) * (1.0 + (double)l2 * dln2));
}
__device__ __forceinline__ double fric_pow(double x, double) { return pow(x, 4.0 / 3.0); }

__device__ __forceinline__ int xcd_block(int i, int G, int remap) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous slice of the logical index space so neighbour gathers stay in
    // that XCD's L2.
    if (!remap || G < 16) return i;
    int x = i & 7, j = i >> 3, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + j;
}

// DPP lane moves within 16-lane rows (gfx9 dpp_ctrl encodings).
enum DppCtrl {
    kDppQuadXor1 = 0xB1,      // quad_perm [1,0,3,2]
    kDppQuadXor2 = 0x4E,      // quad_perm [2,3,0,1]
    kDppRowHalfMirror = 0x141,
    kDppRowMirror = 0x140,
};

template <int CTRL>
__device__ __forceinline__ float dpp_move(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum (MIN = false) or minimum over each 16-lane row; every lane of the row
// ends with the same value (each step combines a lane with its mirror, so both
// sides add the same two operands).
template <bool MIN, typename T>
__device__ __forceinline__ T row_reduce(T v) {
    auto op = [](T a, T b) { return MIN ? (b < a ? b : a) : a + b; };
    v = op(v, dpp_move<kDppQuadXor1>(v));
    v = op(v, dpp_move<kDppQuadXor2>(v));
    v = op(v, dpp_move<kDppRowHalfMirror>(v));
    v = op(v, dpp_move<kDppRowMirror>(v));
    return v;
}

// Wave sums of N <= 4 values at the cost of about one: the value set is split
// between the wave halves (exchange with lane ^ 32) and rows (lane ^ 16) as it
// is reduced, then each row is summed with DPP moves.  Afterwards the lanes of
// row group k (64 / NP lanes, NP = N rounded up to a power of two) hold the sum
// of value k.  Fixed order: deterministic.
template <int N, typename T>
__device__ __forceinline__ T wave_sum_split(const T (&v)[N], int lane) {
    static_assert(N >= 1 && N <= 4, "wave_sum_split takes 1 to 4 values");
    constexpr int NP = N == 1 ? 1 : (N == 2 ? 2 : 4);
    T a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = k < N ? v[k] : T(0);
    const bool up32 = (lane & 32) != 0, up16 = (lane & 16) != 0;
    T x;
    if (NP == 4) {
        // lower half keeps values 0, 1; upper half keeps 2, 3
        const T s0 = up32 ? a[0] : a[2], s1 = up32 ? a[1] : a[3];
        const T o0 = up32 ? a[2] : a[0], o1 = up32 ? a[3] : a[1];
        const T x0 = o0 + __shfl_xor(s0, 32, 64);
        const T x1 = o1 + __shfl_xor(s1, 32, 64);
        // the first row of each half keeps x0, the second x1
        const T s = up16 ? x0 : x1, o = up16 ? x1 : x0;
        x = o + __shfl_xor(s, 16, 64);
    } else if (NP == 2) {
        const T s = up32 ? a[0] : a[1], o = up32 ? a[1] : a[0];
        x = o + __shfl_xor(s, 32, 64);
        x = x + __shfl_xor(x, 16, 64);
    } else {
        x = a[0] + __shfl_xor(a[0], 32, 64);
        x = x + __shfl_xor(x, 16, 64);
    }
    return row_reduce<false>(x);
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    T w = __shfl_xor(v, 32, 64);
    v = w < v ? w : v;
    w = __shfl_xor(v, 16, 64);
    v = w < v ? w : v;
    return row_reduce<true>(v);
}

// Compiler-only memory barrier: loads written above it are issued above it.
// Kernels use it to start their gathers before testing the member's control
// block (a test the compiler would otherwise place first, sinking the loads
// into the taken path).  It emits no instruction and waits for nothing.
#define FH_ISSUE_LOADS() asm volatile("" ::: "memory")
// The same for loads through __restrict__ kernel arguments, which the compiler
// would otherwise move across the barrier: naming the pointers makes the
// barrier a possible writer of their memory.
#define FH_ISSUE_LOADS_OF(...) asm volatile("" ::__VA_ARGS__ : "memory")

// Deterministic block sum of N values (fixed butterfly + fixed wave order).
// Result valid in thread 0.  lds must hold N*16 entries.
template <int N, typename T>
__device__ __forceinline__ void block_sum(T (&v)[N], T* lds) {
    static_assert(N >= 1 && N <= 4, "block_sum takes 1 to 4 values");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    constexpr int STRIDE = N == 1 ? 64 : (N == 2 ? 32 : 16);
    const T x = wave_sum_split<N>(v, lane);
    if ((lane & (STRIDE - 1)) == 0 && lane / STRIDE < N) lds[(lane / STRIDE) * 16 + w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            T s = lds[k * 16];
            for (int i = 1; i < nw; ++i) s += lds[k * 16 + i];
            v[k] = s;
        }
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ T block_min(T v, T* lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    v = wave_min(v);
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = lds[0];
        for (int i = 1; i < nw; ++i) s = lds[i] < s ? lds[i] : s;
        v = s;
    }
    __syncthreads();
    return v;
}

// Adds one block's partials to the block's own accumulation slot.  Only this
// block ever writes the slot and launches are stream-ordered, so the slots
// need no atomics; k_fold sums them into the FP64 ledger in a fixed order.
template <int N>
__device__ __forceinline__ void accumulate_slot(double* acc, int slot, const double (&v)[N]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) acc[(size_t)slot * N + k] += v[k];
    }
}

// accumulate_slot with the slot's old value loaded at the start of the kernel,
// so the final read-modify-write costs no memory round trip (same sum: old + v).
template <int N>
struct SlotAcc {
    double old[N];
    __device__ __forceinline__ void prefetch(const double* acc, int slot) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < N; ++k) old[k] = acc[(size_t)slot * N + k];
        }
    }
    __device__ __forceinline__ void commit(double* acc, int slot, const double (&v)[N]) const {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < N; ++k) acc[(size_t)slot * N + k] = old[k] + v[k];
        }
    }
};

}  // namespace fh
