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

// Completion counters of the last-block reductions (one per reducing kernel).
enum Counter { C_BEGIN = 0, C_FLUX, C_UPDATE, C_INLET, C_NODE, C_RETURN, C_EXTERNAL, C_FINITE,
               C_COUNT = 16 };

// Step control block, resident in device memory.  Written only by the last
// block of k_begin (clock, dt, forcing lookups, flags), by block 0 of
// k_update (eval_next) and by the last block of each ledger-committing kernel
// (diag/ledger terms); every other access is a read.
struct Ctl {
    double t;              // FP64 simulated time (start of the next step)
    double t_step;         // start time of the current step
    double target;         // stop when t >= target
    double final_t;        // event end time (dt cap final_t - t)
    double dt;             // dt of the current step, as applied (== (double)dt_real)
    double dt_real;        // dt of the current step in state precision (stored as double)
    double dt_wet;         // dt of the previous executed step: pending wet-time increment
    double cached_cfl;     // last evaluated CFL dt (state precision value)
    double rain_rate;      // rain rate of the current step (state precision value)
    double ext_frac;       // external inflow interpolation fraction at t + dt/2
    int64_t step;          // global step counter
    int64_t interval_step; // steps taken in the current advance call
    int64_t max_steps;     // step budget of the current advance call
    int64_t cfl_evals;
    int32_t ext_lower, ext_upper;
    int32_t active;        // the current step executes
    int32_t pending_wet;   // a wet-time increment is pending (applied by the next update / finalize)
    int32_t eval_next;     // the next step evaluates the CFL condition
    int32_t override_dt;   // host-driven single step: dt and rain given by the host
    double override_dt_value;
    double override_rain;
    int32_t nonfinite;     // count of non-finite values found by k_finite
    int32_t pad0;
    // forcing-table search hints: the table position found at an earlier time of
    // this clock (reset whenever the clock or the forcing is set); used only when
    // the table's end times are sorted, which the host checks
    int32_t rain_hint, ext_hint, rain_sorted, ext_sorted;
    // the rain interval at rain_hint (sorted tables): its end (the next end time
    // after t; -inf forces a lookup), start and rate -- a step whose t is still
    // below that end reads its rate here instead of in the table
    double rain_next_end, rain_cur_start, rain_cur_rate;
    double cfl_min;        // result of the last CFL reduction
    unsigned long long cfl_bits;  // running min of the CFL reduction (bits of a non-negative double)
    double diag[kLedgerTerms];    // this step's ledger increments
    double ledger[kLedgerTerms];  // cumulative FP64 ledger
    unsigned counters[C_COUNT];
};

// Block -> member mapping of one index space (cells, edges, nodes, ...).  Each
// member's items start on a tile boundary (tile = blockDim.x * items per
// thread), so a block never straddles two members.
struct Seg {
    const int* blk_member;   // [blocks]   member of each (logical) block
    const int* mem_blk0;     // [members]  first block of the member
    const int* mem_nblk;     // [members]  number of blocks of the member
    const int* mem_end;      // [members]  exclusive end of the member's real items
};

template <typename real>
struct KP {                 // kernel constants, pre-rounded to the state precision
    real g, eps, half_g, two_g, cfl, max_speed, tiny, capture_q, pow43, reuse, slot_coef;
    double max_dt, dt_floor, fixed_dt;
    int interval;
    int remap;
};

// NaN-propagating max/min, like numpy.maximum / jax.numpy.maximum (fmax/fmin
// would drop a NaN operand and hide a blown-up state from the finite checks).
template <typename T>
__device__ __forceinline__ T vmax(T a, T b) { return (a != a || a > b) ? a : b; }
template <typename T>
__device__ __forceinline__ T vmin(T a, T b) { return (a != a || a < b) ? a : b; }

__device__ __forceinline__ int xcd_block(int i, int G, int remap) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous slice of the logical index space so neighbour gathers stay in
    // that XCD's L2.
    if (!remap || G < 16) return i;
    int x = i & 7, j = i >> 3, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + j;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, 64); v = w < v ? w : v; }
    return v;
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
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) lds[k * 16 + w] = v[k];
    }
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

// Last-block-done: every block of the reduction publishes its partials, then
// the block that arrives last (grid-wide counter) sums all partials in a fixed
// order and commits them.  Returns true in the last block (all threads).
//
// The partials are written and read with agent-scope atomic accesses, which are
// coherent across the XCDs' L2 caches, so no cache-wide fence is needed: thread
// 0 only waits until its own partial stores have completed before it bumps the
// counter.  (An agent-scope __threadfence here would write back the whole L2 of
// the XCD once per block.)
__device__ __forceinline__ bool arrive_last(unsigned* counter, unsigned nblocks) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_s_waitcnt(0);
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (prev == nblocks - 1);
    }
    __syncthreads();
    return s_last != 0;
}

// Sum `nparts` partial records of N doubles (stride N) with the whole block, in a
// fixed order.  Result in out[] of thread 0.
template <int N>
__device__ __forceinline__ void sum_partials(const double* part, int nparts, double (&out)[N],
                                             double* lds) {
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            out[k] += __hip_atomic_load(part + (size_t)i * N + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<N>(out, lds);
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

// Publish one block's partial record (agent-scope stores: visible to the last
// block wherever it runs, across XCD L2s).
template <int N>
__device__ __forceinline__ void publish_partials(double* part, int index, const double (&v)[N]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            __hip_atomic_store(part + (size_t)index * N + k, v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace fh
