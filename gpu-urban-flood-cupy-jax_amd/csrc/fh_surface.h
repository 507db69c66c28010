// Surface step kernels: CFL/dt control, per-edge HLL flux, per-cell CSR-gather
// update with friction, sources and velocity clamp fused.
//
// Every kernel serves a batch of members (independent scenarios, each with its
// own clock, dt, forcing and ledger): the index spaces are concatenated member
// by member on tile boundaries and a Seg table maps each block to its member.
// A single scenario is the one-member case.
//
// Semantics follow hybrid_flood/jax_solver/v1_parity.py (paths relative to
// HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/) operation by
// operation, so float32 results round like the reference's float32 run:
//   _hll :110-132, _hydrostatic_flux_pair :135-150, stable_timestep :153-158,
//   surface_flux_step :161-214, apply_rainfall_horton :217-242,
//   clamp_velocity :245-256; dt policy parity_runner.py:124-135.
#pragma once
#include "fh_device.h"

namespace fh {

// ---------------------------------------------------------------------------
// edge encodings (set up on the host from edge_right / boundary_code / active)
//   right >= 0 : both cells active, interior edge
//   right == -1: ghost copies the left state (transmissive / unsupported codes)
//   right == -2: reflective ghost (codes 0 and 2, or an inactive right cell)
//   right == -3: outflow-only boundary (code 4): copy, inflow blocked
// Edges whose left cell is inactive contribute exactly nothing and are dropped,
// except code-3 boundary edges (kGhostDeadStage / kGhostDeadOther below).
// ---------------------------------------------------------------------------
constexpr int kGhostCopy = -1;
constexpr int kGhostReflect = -2;
constexpr int kGhostOutflow = -3;
// stage boundary (code 3) of a member with stage values: the V1 Dirichlet stage
// ghost (triangular_backend.py:1294-1327: depth max(stage - z_left, 0), zero
// momentum, the left cell's bed), whose depth the host precomputes per edge.
// A wall in the local-inertial step, which lets water out only at codes 4 and 1.
constexpr int kGhostStage = -4;
// code 3 without stage values (a copy ghost, as in v1_parity.py, which has no
// stage values) and any code other than 0-4: a copy ghost in the full step, a
// wall in the local-inertial step
constexpr int kGhostOther = -5;
// A code-3 boundary edge whose cell is inactive is kept, unlike the other edges
// of inactive cells: the V1 step computes its flux from a zero left state, so a
// stage above the cell's bed flows in, is counted as boundary inflow and is then
// removed with the inactive cell's water (v1_parity.py:170-214,
// triangular_backend.py:1294-1327).  Without stage values the flux is zero.
// Both kinds are walls in the local-inertial step.
constexpr int kGhostDeadStage = -6;
constexpr int kGhostDeadOther = -7;

// cell -> edge table entry: (edge_position << 2) | (boundary << 1) | is_right_side, -1 none,
// -2 in slot 0 marks an inactive (or padding) cell, whose slots 1 and 2 hold its
// kept code-3 boundary edges, if any.
constexpr int kCellInactive = -2;

template <typename real>
struct EdgeFlux { real m, fuL, fvL, fuR, fvR; };

// Hydrostatic reconstruction + HLL for one edge, returning flux * length for the
// left side (m, fuL, fvL) and the right side momentum (fuR, fvR).
template <typename real>
__device__ __forceinline__ EdgeFlux<real> edge_flux(real h_l, real hu_l, real hv_l, real z_l,
                                                    real h_r, real hu_r, real hv_r, real z_r,
                                                    real nx, real ny, real len, bool block_inflow,
                                                    const KP<real>& p) {
    const real zero = real(0), eps = p.eps;
    // _hydrostatic_flux_pair
    h_l = vmax(h_l, zero);
    h_r = vmax(h_r, zero);
    const real z_star = vmax(z_l, z_r);
    const real hs_l = vmax(h_l + z_l - z_star, zero);
    const real hs_r = vmax(h_r + z_r - z_star, zero);
    const bool wl0 = h_l > eps, wr0 = h_r > eps;
    const Divisor<real> dl0(vmax(h_l, eps)), dr0(vmax(h_r, eps));
    const real u_l0 = wl0 ? dl0.div(hu_l) : zero;
    const real v_l0 = wl0 ? dl0.div(hv_l) : zero;
    const real u_r0 = wr0 ? dr0.div(hu_r) : zero;
    const real v_r0 = wr0 ? dr0.div(hv_r) : zero;
    const real qhu_l = u_l0 * hs_l, qhv_l = v_l0 * hs_l;
    const real qhu_r = u_r0 * hs_r, qhv_r = v_r0 * hs_r;
    // _hll on the reconstructed states
    const real a_l = vmax(hs_l, zero), a_r = vmax(hs_r, zero);
    const bool wl = a_l > eps, wr = a_r > eps;
    real fh = zero, fhu = zero, fhv = zero;
    if (wl || wr) {
        const Divisor<real> dl(vmax(a_l, eps)), dr(vmax(a_r, eps));
        const real ul = wl ? dl.div(qhu_l) : zero;
        const real vl = wl ? dl.div(qhv_l) : zero;
        const real ur = wr ? dr.div(qhu_r) : zero;
        const real vr = wr ? dr.div(qhv_r) : zero;
        const real un_l = ul * nx + vl * ny;
        const real un_r = ur * nx + vr * ny;
        const real c_l = sqrt(p.g * a_l), c_r = sqrt(p.g * a_r);
        const real p_l = p.half_g * (a_l * a_l), p_r = p.half_g * (a_r * a_r);
        const real fl0 = a_l * un_l, fl1 = qhu_l * un_l + p_l * nx, fl2 = qhv_l * un_l + p_l * ny;
        const real fr0 = a_r * un_r, fr1 = qhu_r * un_r + p_r * nx, fr2 = qhv_r * un_r + p_r * ny;
        const real s_l = vmin(un_l - c_l, un_r - c_r);
        const real s_r = vmax(un_l + c_l, un_r + c_r);
        if (s_l >= zero) {
            fh = fl0; fhu = fl1; fhv = fl2;
        } else if (s_r <= zero) {
            fh = fr0; fhu = fr1; fhv = fr2;
        } else {
            const real ds = s_r - s_l;
            // s_l < 0 < s_r here, so the divisor is positive
            const Divisor<real> denom(fabs(ds) > p.tiny ? ds : real(1));
            const real slsr = s_l * s_r;
            fh = denom.div(s_r * fl0 - s_l * fr0 + slsr * (a_r - a_l));
            fhu = denom.div(s_r * fl1 - s_l * fr1 + slsr * (qhu_r - qhu_l));
            fhv = denom.div(s_r * fl2 - s_l * fr2 + slsr * (qhv_r - qhv_l));
        }
    }
    const real pc_l = p.half_g * (h_l * h_l - hs_l * hs_l);
    const real pc_r = p.half_g * (h_r * h_r - hs_r * hs_r);
    real Lh = fh + zero, Lu = fhu + pc_l * nx, Lv = fhv + pc_l * ny;
    const real Ru = fhu + pc_r * nx, Rv = fhv + pc_r * ny;
    if (block_inflow && Lh < zero) { Lh = zero; Lu = zero; Lv = zero; }
    EdgeFlux<real> out;
    out.m = Lh * len;
    out.fuL = Lu * len;
    out.fvL = Lv * len;
    out.fuR = Ru * len;
    out.fvR = Rv * len;
    return out;
}

// Per-member forcing tables of the step plan.
struct ForcingTables {
    const int* rain_off;     // [M] offset of the member's rain intervals
    const int* rain_n;       // [M] number of rain intervals
    const int* ext_off;      // [M] offset of the member's external time table
    const int* ext_n;        // [M] number of external times (0: no external inflow)
    const float* rain_start;
    const float* rain_end;
    const float* rain_rate;
    const float* ext_times;
};

// The CFL time of one cell (v1_parity.py:153-158): cfl * Lc / (|u| + sqrt(g h)),
// +inf where that is not finite (so the minimum skips it).
template <typename real>
__device__ __forceinline__ real cell_cfl(const KP<real>& p, const typename V4<real>::type& s, real lc) {
    const real hs = vmax(s.x, p.eps);
    const Divisor<real> dh(hs);
    const real u = dh.div(s.y), v = dh.div(s.z);
    real speed = sqrt(u * u + v * v);
    speed = speed + sqrt(p.g * hs);
    const real local = Divisor<real>(vmax(speed, p.tiny)).div(p.cfl * lc);
    return isfinite(local) ? local : real(INFINITY);
}

// The local-inertial CFL time of one cell (equations.py:181-207): gravity-wave
// speed only, cfl * Lc / max(sqrt(g max(h, min depth)), 1e-12), +inf where not
// finite.
template <typename real>
__device__ __forceinline__ real cell_cfl_li(const KP<real>& p, const typename V4<real>::type& s, real lc) {
    const real hs = vmax(s.x, p.eps);
    const real c = sqrt(p.g * hs);
    const real local = Divisor<real>(vmax(c, p.tiny)).div(p.cfl * lc);
    return isfinite(local) ? local : real(INFINITY);
}

template <typename real, bool LI>
__device__ __forceinline__ real cell_time(const KP<real>& p, const typename V4<real>::type& s, real lc) {
    if (LI) return cell_cfl_li(p, s, lc);
    return cell_cfl(p, s, lc);
}

// Block minimum, then one atomic min per block on the bits of a non-negative
// double (which order like the values).
// Block minimum of v into the CFL slot of block `blk` (member m, step parity).
__device__ __forceinline__ void commit_cfl_min(unsigned long long* cfl, int m, int parity, int blk, double v,
                                               double* lds) {
    v = block_min(v, lds);
    if (threadIdx.x == 0)
        __hip_atomic_fetch_min(cfl_slot(cfl, m, parity, blk % kCflSlots), (unsigned long long)__double_as_longlong(v),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The step's CFL minimum (member m, parity): each lane of the wave reads one of
// the kCflSlots (= 64, the wave width) slots, a butterfly takes their minimum,
// and the result is made wave-uniform.  The values are non-negative or +inf.
__device__ __forceinline__ double cfl_slots_min(unsigned long long* cfl, int m, int parity) {
    static_assert(kCflSlots == 64, "one slot per lane of a 64-wide wave");
    double v = __longlong_as_double((long long)*cfl_slot(cfl, m, parity, (int)(threadIdx.x & 63)));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Rain rate at time t (parity_runner.py:96-102): rate[lo] when lo =
// searchsorted(end, t, side="right") < n and t >= start[lo], else 0.  The cache
// holds the interval of an earlier t of the same clock; with sorted end times the
// position only grows with t, so the lookup scans forward from there, and is
// skipped while t stays below the cached end.
__device__ __forceinline__ double rain_rate_at(const ForcingTables& ft, int m, double t, bool sorted,
                                               RainCache& rc) {
    if (!(t < rc.next_end)) {
        const int n = ft.rain_n[m];
        const int off = ft.rain_off[m];
        const float* re = ft.rain_end + off;
        int lo = 0;
        if (sorted) {
            lo = rc.hint < n ? rc.hint : n;
            while (lo < n && (double)re[lo] <= t) ++lo;
        } else {
            int hi = n;
            while (lo < hi) { const int mid = (lo + hi) >> 1; if ((double)re[mid] <= t) lo = mid + 1; else hi = mid; }
        }
        rc.hint = lo;
        rc.start = lo < n ? (double)ft.rain_start[off + lo] : INFINITY;
        rc.rate = lo < n ? (double)ft.rain_rate[off + lo] : 0.0;
        rc.next_end = sorted ? (lo < n ? (double)re[lo] : INFINITY) : -INFINITY;
    }
    return t >= rc.start ? rc.rate : 0.0;
}

// The plan of one step of one member: the dt policy of parity_runner.py:124-135
// and the rain rate of :96-102, from the start-of-step clock.  Every block of
// the update kernel computes it from the same inputs, so all agree bit for bit.
template <typename real>
struct StepPlan {
    bool active;
    bool eval_used;          // the CFL minimum of this step was evaluated and used
    bool cache_evaluated;    // the policy path that updates cached_cfl / cfl_evals
    double dt;               // (double)(real) dt
    double rate;             // rain rate
    double dt_wet;           // pending wet-time increment of the previous step
    double evaluated;        // the CFL dt this step evaluated or reused
    RainCache rain;          // rain cache after the lookup at t
};

// Whether the member's step executes, as plan_step decides it.  The kernels of a
// batch (GATE: the context holds more than one member) test it, or the clock's
// copy of it, before their first load: a member that has reached its output
// time while others step on then costs its blocks' dispatch, not its bytes.
// Every member of a batch stops at its own time, and in the members' last steps
// most blocks belong to idle members.  The plain solver's kernels (GATE false)
// load ahead of the test instead, so their gathers start one memory round trip
// earlier.
__device__ __forceinline__ bool step_runs(const Ctl* ctl, int parity) {
    const Clock& c = ctl->clk[parity];
    return (c.t < ctl->target) && (c.interval_step < ctl->max_steps);
}

template <typename real>
__device__ __forceinline__ StepPlan<real> plan_step(const Ctl* ctl, const Clock& c, const KP<real>& p,
                                                    const ForcingTables& ft, int m, double gmin) {
    StepPlan<real> pl{};
    const double t = c.t;
    pl.active = (t < ctl->target) && (c.interval_step < ctl->max_steps);
    pl.rain = c.rain;
    if (!pl.active) return pl;
    pl.dt_wet = c.pending_wet ? c.dt_prev : 0.0;
    double dt;
    double rate;
    if (ctl->override_dt) {
        dt = (double)(real)ctl->override_dt_value;
        rate = ctl->override_rain;
    } else {
        if (p.fixed_dt > 0.0) {
            dt = p.fixed_dt;
        } else {
            pl.eval_used = c.eval != 0;
            pl.evaluated = pl.eval_used ? gmin : c.cached_cfl;
            const double candidate = pl.eval_used ? pl.evaluated : (double)((real)pl.evaluated * p.reuse);
            dt = fmax(fmin(fmin(candidate, p.max_dt), ctl->final_t - t), p.dt_floor);
            pl.cache_evaluated = true;
        }
        rate = rain_rate_at(ft, m, t, ctl->rain_sorted != 0, pl.rain);
    }
    pl.dt = (double)(real)dt;
    pl.rate = (double)(real)rate;
    return pl;
}

// Written by one thread of the member's update, after its cells: the step
// outputs (dt, rain rate, the external inflow interpolation of
// parity_runner.py:105-115 at t + dt/2, for the drainage kernels) and the clock
// of the next step (clk[1 - parity]) with its rain interval already looked up,
// and a fresh CFL minimum slot for that step.
template <typename real>
__device__ __forceinline__ void commit_plan(Ctl* ctl, int parity, const Clock& c, const StepPlan<real>& pl,
                                            const KP<real>& p, const ForcingTables& ft, int m, double gmin) {
    Clock n = c;
    ctl->active = pl.active ? 1 : 0;
    if (pl.active) {
        ctl->dt = pl.dt;
        ctl->dt_real = pl.dt;
        ctl->rain_rate = pl.rate;
        ctl->t_step = c.t;
        const int n_ext = ft.ext_n[m];
        if (n_ext > 0) {
            const float* et = ft.ext_times + ft.ext_off[m];
            const double te = c.t + 0.5 * pl.dt;
            int lo = 0;
            if (ctl->ext_sorted) {     // te = t + dt/2 never decreases along a clock
                lo = c.ext_hint < n_ext ? c.ext_hint : n_ext;
                while (lo < n_ext && (double)et[lo] <= te) ++lo;
                n.ext_hint = lo;
            } else {
                int hi = n_ext;
                while (lo < hi) { const int mid = (lo + hi) >> 1; if ((double)et[mid] <= te) lo = mid + 1; else hi = mid; }
            }
            const int upper = lo;
            int lower = upper - 1; lower = lower < 0 ? 0 : (lower > n_ext - 1 ? n_ext - 1 : lower);
            const int us = upper > n_ext - 1 ? n_ext - 1 : upper;
            const double den = fmax((double)et[us] - (double)et[lower], 1.0e-12);
            double frac = (te - (double)et[lower]) / den;
            frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
            ctl->ext_lower = lower;
            ctl->ext_upper = us;
            ctl->ext_frac = frac;
        }
        if (pl.eval_used) ctl->cfl_min = gmin;
        n.t = c.t + pl.dt;
        n.step = c.step + 1;
        n.interval_step = c.interval_step + 1;
        if (pl.cache_evaluated) {
            n.cached_cfl = pl.evaluated;
            if (pl.eval_used) n.cfl_evals = c.cfl_evals + 1;
        }
        n.dt_prev = pl.dt;
        n.pending_wet = 1;
        n.rain = pl.rain;
        if (ctl->rain_sorted) (void)rain_rate_at(ft, m, n.t, true, n.rain);
        n.eval = (p.fixed_dt <= 0.0 && ctl->override_dt == 0 && (n.step % p.interval) == 0) ? 1 : 0;
        n.active = (n.t < ctl->target) && (n.interval_step < ctl->max_steps);
    }
    ctl->clk[1 - parity] = n;
    for (int k = 0; k < kCflSlots; ++k)
        *cfl_slot(p.cfl_slots, m, 1 - parity, k) = (unsigned long long)__double_as_longlong(INFINITY);
    ctl->li_cfl_step[1 - parity] = -1;
}

// ---------------------------------------------------------------------------
// k_cfl: CFL minimum over every cell of each member (stable_timestep,
// v1_parity.py:153-158); wave shuffle -> LDS -> one atomic min per block into
// the member's slots, which k_cfl_final reduces into cfl_min.
// ---------------------------------------------------------------------------
struct CellRanges {
    Seg seg;                 // k_cfl blocks of each member
    const int* cell0;        // [M] first cell of the member
    const int* cell1;        // [M] end of the member's real cells
};

template <typename real, bool LI>
__global__ void __launch_bounds__(256) k_cfl(KP<real> p, int parity,
        const typename V4<real>::type* __restrict__ state, const real* __restrict__ lc, CellRanges cr) {
    __shared__ double lds[16 * 2];
    const int b = blockIdx.x;
    const int m = cr.seg.blk_member[b];
    const int blk0 = cr.seg.mem_blk0[m], nblk = cr.seg.mem_nblk[m];
    real mn = real(INFINITY);
    const int c1 = cr.cell1[m];
    for (int c = cr.cell0[m] + (b - blk0) * blockDim.x + threadIdx.x; c < c1; c += nblk * blockDim.x) {
        const real local = cell_time<real, LI>(p, state[c], lc[c]);
        if (local < mn) mn = local;
    }
    // the next step's slots (parity) hold +inf between steps; k_cfl_final reads
    // them after this launch has completed and restores +inf
    commit_cfl_min(p.cfl_slots, m, parity, b - blk0, (double)mn, lds);
}

// One 64-lane wave per member, launched after k_cfl on the same stream: the
// launch boundary orders every block's slot update before these reads, so the
// reduction needs no grid-wide completion counter.
__global__ void __launch_bounds__(64) k_cfl_final(Ctl* __restrict__ ctls, unsigned long long* __restrict__ cfl,
                                                  int parity) {
    const int m = blockIdx.x;
    const double v = cfl_slots_min(cfl, m, parity);
    *cfl_slot(cfl, m, parity, (int)threadIdx.x) = (unsigned long long)__double_as_longlong(INFINITY);
    if (threadIdx.x == 0) ctls[m].cfl_min = v;
}

// Edge index space: Seg, plus the blocks after the edge blocks that cover each
// member's cells which no edge reaches (inactive cells, cells whose every edge
// was dropped), so a CFL evaluation in the flux kernel sees every cell.
struct EdgeSeg {
    Seg seg;
    int edge_blocks;         // blocks of the edge space; the orphan blocks follow
    Seg oseg;                // orphan-cell blocks of each member
    const int* ocells;       // [orphans] cell of each orphan slot
};

// ---------------------------------------------------------------------------
// k_flux: EPT edges per thread (tile = blockDim * EPT, coalesced per k).  Writes
// the edge record rec[e] = (m, fuL, fvL, fuR), recR[e] = fvR (all * length).  On
// a step that evaluates the CFL condition it also reduces the CFL time of both
// cells of every edge (and of the orphan cells) into the step's minimum slot:
// every cell counts at least once, and a cell counted twice gives the same value.
// SH: shared-mesh batch; the edge tables hold member-local cells and are indexed
// by the member-local edge.  stage_h: the stage ghost depth of each edge position
// (read at stage edges only).
// ---------------------------------------------------------------------------
template <typename real, int EPT, bool SH, bool GATE>
__global__ void __launch_bounds__(256) k_flux(Ctl* __restrict__ ctls, KP<real> p, EdgeSeg es, int parity,
        const int* __restrict__ el, const int* __restrict__ er, const real* __restrict__ enx,
        const real* __restrict__ eny, const real* __restrict__ elen,
        const typename V4<real>::type* __restrict__ state, const real* __restrict__ lc,
        typename V4<real>::type* __restrict__ rec, real* __restrict__ recR, const real* __restrict__ stage_h) {
    typedef typename V4<real>::type V;
    __shared__ double lds[16];
    if ((int)blockIdx.x >= es.edge_blocks) {
        // orphan cells: CFL only
        const int ob = blockIdx.x - es.edge_blocks;
        const int m = es.oseg.blk_member[ob];
        Ctl* ctl = ctls + m;
        const Clock& ck = ctl->clk[parity];
        if (!ck.active || !ck.eval || ctl->override_dt || p.fixed_dt > 0.0) return;
        const int o = ob * (int)blockDim.x + threadIdx.x;
        real mn = real(INFINITY);
        if (o < es.oseg.mem_end[m]) {
            const int c = es.ocells[o];
            mn = cell_cfl(p, state[c], lc[c]);
        }
        commit_cfl_min(p.cfl_slots, m, parity, ob, (double)mn, lds);
        return;
    }
    const int i = xcd_block(blockIdx.x, es.edge_blocks, p.remap);
    int b, m, g0 = 0, cb = 0;
    if constexpr (SH) {
        b = shared_block(es.seg, i, m);
        g0 = m * es.seg.sh_items;
        cb = m * es.seg.sh_link;
    } else {
        b = i;
        m = es.seg.blk_member[b];
    }
    if (GATE && !ctls[m].clk[parity].active) return;
    // every load ahead of the control block: the grid covers the padded edge
    // space exactly, padding edges point at cell 0 as copy ghosts
    int lft[EPT], rgt[EPT];
    real nxs[EPT], nys[EPT], lens[EPT];
    V sls[EPT], srs[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int g = (b * EPT + k) * (int)blockDim.x + threadIdx.x - g0;
        lft[k] = el[g] + cb;
        const int r = er[g];
        rgt[k] = SH && r >= 0 ? r + cb : r;
        nxs[k] = enx[g]; nys[k] = eny[g]; lens[k] = elen[g];
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        sls[k] = state[lft[k]];
        srs[k] = state[rgt[k] >= 0 ? rgt[k] : lft[k]];   // ghost edges: unused
    }
    const int end = es.seg.mem_end[m];
    FH_ISSUE_LOADS_OF("s"(el), "s"(er), "s"(enx), "s"(eny), "s"(elen), "s"(state));
    Ctl* ctl = ctls + m;
    const Clock& ck = ctl->clk[parity];
    if (!ck.active) return;
    const bool eval = ck.eval && !ctl->override_dt && !(p.fixed_dt > 0.0);
    real mn = real(INFINITY);
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = (b * EPT + k) * (int)blockDim.x + threadIdx.x;
        if (e >= end) continue;
        const int r = rgt[k];
        const real nx = nxs[k], ny = nys[k], len = lens[k];
        V sl = sls[k];
        real h_r, hu_r, hv_r, z_r;
        if (r >= 0) {
            const V sr = srs[k];
            h_r = sr.x; hu_r = sr.y; hv_r = sr.z; z_r = sr.w;
        } else if (r == kGhostReflect) {
            const real nm = sl.y * nx + sl.z * ny;
            h_r = sl.x; hu_r = sl.y - real(2) * nm * nx; hv_r = sl.z - real(2) * nm * ny; z_r = sl.w;
        } else if (r == kGhostStage || r == kGhostDeadStage) {
            h_r = stage_h[e]; hu_r = real(0); hv_r = real(0); z_r = sl.w;
        } else {
            h_r = sl.x; hu_r = sl.y; hv_r = sl.z; z_r = sl.w;
        }
        if (r <= kGhostDeadStage) {   // inactive cell: zero left state (and copy ghost)
            sl.x = real(0); sl.y = real(0); sl.z = real(0);
            if (r == kGhostDeadOther) { h_r = real(0); hu_r = real(0); hv_r = real(0); }
        }
        const EdgeFlux<real> f = edge_flux<real>(sl.x, sl.y, sl.z, sl.w, h_r, hu_r, hv_r, z_r, nx, ny, len,
                                                 r == kGhostOutflow, p);
        V o;
        o.x = f.m; o.y = f.fuL; o.z = f.fvL; o.w = f.fuR;
        rec[e] = o;
        if (r >= 0) recR[e] = f.fvR;
        if (eval) {
            const real cl = cell_cfl(p, sls[k], lc[lft[k]]);
            mn = cl < mn ? cl : mn;
            if (r >= 0) {
                const real cr = cell_cfl(p, srs[k], lc[r]);
                mn = cr < mn ? cr : mn;
            }
        }
    }
    if (eval) commit_cfl_min(p.cfl_slots, m, parity, b, (double)mn, lds);
}

// ---------------------------------------------------------------------------
// Local-inertial surface mode (triangular_backend.py:1181-1291), three launches
// after the step's CFL pass: k_li_flux (edge discharges), k_li_donor (donor
// limiting), k_update<..., LI> (continuity, boundary capture, momentum, then the
// shared rain/Horton/clamp).  The discharge needs the step's dt, so the CFL
// minimum of an evaluating step comes from k_li_cfl, launched first.
// ---------------------------------------------------------------------------

// The step's local-inertial CFL minimum over every cell of each member, on the
// steps that evaluate it (every other step returns at once): the same slot the
// full step's flux kernel fills.
template <typename real>
__global__ void __launch_bounds__(256) k_li_cfl(Ctl* __restrict__ ctls, KP<real> p, int parity,
        const typename V4<real>::type* __restrict__ state, const real* __restrict__ lc, CellRanges cr) {
    __shared__ double lds[16 * 2];
    const int b = blockIdx.x;
    const int m = cr.seg.blk_member[b];
    Ctl* ctl = ctls + m;
    const Clock& ck = ctl->clk[parity];
    if (!ck.active || !ck.eval || ctl->override_dt || p.fixed_dt > 0.0) return;
    const int blk0 = cr.seg.mem_blk0[m], nblk = cr.seg.mem_nblk[m];
    real mn = real(INFINITY);
    const int c1 = cr.cell1[m];
    for (int c = cr.cell0[m] + (b - blk0) * blockDim.x + threadIdx.x; c < c1; c += nblk * blockDim.x) {
        const real local = cell_cfl_li(p, state[c], lc[c]);
        if (local < mn) mn = local;
    }
    commit_cfl_min(p.cfl_slots, m, parity, b - blk0, (double)mn, lds);
    if (b == blk0 && threadIdx.x == 0) ctl->li_cfl_step[parity] = ck.step;
}

// One edge's local-inertial discharge (triangular_backend.py:1196-1248 with
// local_inertial.py:20-48): the LISFLOOD-FP update of the face discharge from
// the water-surface slope between the two cells, friction semi-implicit;
// outflow-only at routable boundary edges (codes 4 and 1), zero at walls
// (codes 0, 2, 3, inactive neighbours).  Writes rec[e] = {Q, Q nx, Q ny} (12
// bytes: the update reads no fourth value) for the update and qv[e] = Q, which
// the donor pass gathers (4 bytes, not 12).
// SH as in k_flux.
template <typename real, bool SH, bool GATE>
__global__ void __launch_bounds__(256) k_li_flux(Ctl* __restrict__ ctls, KP<real> p, EdgeSeg es, int parity,
        ForcingTables ft, const int* __restrict__ el, const int* __restrict__ er, const real* __restrict__ enx,
        const real* __restrict__ eny, const real* __restrict__ elen,
        const typename V4<real>::type* __restrict__ state, const real* __restrict__ lc,
        const real* __restrict__ man, Rec3<real>* __restrict__ rec, real* __restrict__ qv) {
    typedef typename V4<real>::type V;
    const int i = xcd_block(blockIdx.x, es.edge_blocks, p.remap);
    int b, m, g0 = 0, cb = 0;
    if constexpr (SH) {
        b = shared_block(es.seg, i, m);
        g0 = m * es.seg.sh_items;
        cb = m * es.seg.sh_link;
    } else {
        b = i;
        m = es.seg.blk_member[b];
    }
    if (GATE && !step_runs(ctls + m, parity)) return;
    const int e = b * (int)blockDim.x + threadIdx.x;
    const int g = e - g0;
    const int l = el[g] + cb, r0 = er[g];
    const int r = SH && r0 >= 0 ? r0 + cb : r0;
    const int rs = r >= 0 ? r : l;
    const real nx = enx[g], ny = eny[g], len = elen[g];
    const V sl = state[l], sr = state[rs];
    const real lcl = lc[l], lcr = lc[rs], nl = man[l], nr = man[rs];
    const int end = es.seg.mem_end[m];
    FH_ISSUE_LOADS_OF("s"(el), "s"(er), "s"(enx), "s"(eny), "s"(elen), "s"(state), "s"(lc), "s"(man));
    Ctl* ctl = ctls + m;
    const Clock ck = ctl->clk[parity];
    const bool eval = ck.active && ck.eval && !ctl->override_dt && !(p.fixed_dt > 0.0);
    const double gmin = eval ? cfl_slots_min(p.cfl_slots, m, parity) : INFINITY;
    const StepPlan<real> pl = plan_step<real>(ctl, ck, p, ft, m, gmin);
    if (!pl.active) return;
    // an evaluating step whose CFL pass did not run (a graph captured for another
    // step phase) would take dt from an empty minimum: flag it for the host
    if (eval && threadIdx.x == 0 && ctl->li_cfl_step[parity] != ck.step) ctl->li_missed = 1;
    if (e >= end) return;
    const real zero = real(0), half = real(0.5);
    const bool interior = r >= 0;
    const bool outlet = r == kGhostOutflow || r == kGhostCopy;
    real q = zero;
    if (interior || outlet) {
        const real dt = (real)pl.dt;
        const real eta_l = sl.w + sl.x;
        const real eta_r = interior ? sr.w + sr.x : sl.w;
        const real bed_star = interior ? vmax(sl.w, sr.w) : sl.w;
        const real h_flow = vmax(vmax(eta_l, eta_r) - bed_star, zero);
        const real distance = vmax(half * (lcl + (interior ? lcr : lcl)), real(1.0e-6));
        const real slope = Divisor<real>(distance).div(eta_r - eta_l);
        const real nm_l = sl.y * nx + sl.z * ny;
        real nm_face = nm_l;
        real n_face = nl;
        if (interior) {
            const real nm_r = sr.y * nx + sr.z * ny;
            nm_face = half * (nm_l + nm_r);
            n_face = half * (nl + nr);
        }
        const real q_old = nm_face * len;
        const real area = vmax(h_flow * len, real(1.0e-8));
        const real radius = vmax(h_flow, real(1.0e-6));
        const real numerator = q_old - p.g * area * dt * slope;
        const real gdt = (real)(p.g_d * pl.dt);
        const real denominator =
            real(1) + Divisor<real>(fric_pow(radius, p.pow43_dln2) * area).div(gdt * (n_face * n_face) * fabs(q_old));
        q = Divisor<real>(denominator).div(numerator);
        if (!interior) q = vmax(q, zero);
    }
    rec[e] = Rec3<real>{q, q * nx, q * ny};
    qv[e] = q;
}

// Donor limiting of the interior transfers (network_kernels.py:45-67), one
// thread per cell: the cell's requested outflows |Q dt| over the interior edges
// it donates to (from-cell when Q dt >= 0, else to-cell), summed in ascending
// edge order (perm: the cell's entry slots in that order, two bits each, 3 =
// end), give the factor min(1, max(h, 0) A / sum); the cell then writes the
// limited transfer sign * |Q dt| * factor of each edge it donates to (every
// interior edge has exactly one donor) into T[edge].  SH: the cell tables (ce*,
// perm) are shared and indexed by the member-local cell; their edge positions
// are member-local.
template <typename real, bool SH, bool GATE>
__global__ void __launch_bounds__(256) k_li_donor(Ctl* __restrict__ ctls, KP<real> p, Seg seg, int parity,
        ForcingTables ft, const int* __restrict__ ce0, const int* __restrict__ ce1, const int* __restrict__ ce2,
        const unsigned char* __restrict__ perm, const real* __restrict__ qv,
        real* __restrict__ T, const typename V4<real>::type* __restrict__ state, const real* __restrict__ area_) {
    const int i = xcd_block(blockIdx.x, gridDim.x, p.remap);
    int b, m, c0 = 0, eb = 0;
    if constexpr (SH) {
        b = shared_block(seg, i, m);
        c0 = m * seg.sh_items;
        eb = m * seg.sh_link;
    } else {
        b = i;
        m = seg.blk_member[b];
    }
    if (GATE && !step_runs(ctls + m, parity)) return;
    const int c = b * blockDim.x + threadIdx.x;
    const int cl = c - c0;
    const int ent[3] = {ce0[cl], ce1[cl], ce2[cl]};
    real qk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) qk[k] = qv[ent[k] >= 0 ? (ent[k] >> 2) + eb : 0];
    const real h = state[c].x, A = area_[c];
    const unsigned order = perm[cl];
    const int mend = seg.mem_end[m];
    FH_ISSUE_LOADS_OF("s"(ce0), "s"(ce1), "s"(ce2), "s"(perm), "s"(qv), "s"(state), "s"(area_));
    Ctl* ctl = ctls + m;
    const Clock ck = ctl->clk[parity];
    const bool eval = ck.active && ck.eval && !ctl->override_dt && !(p.fixed_dt > 0.0);
    const double gmin = eval ? cfl_slots_min(p.cfl_slots, m, parity) : INFINITY;
    const StepPlan<real> pl = plan_step<real>(ctl, ck, p, ft, m, gmin);
    if (!pl.active || c >= mend || ent[0] == kCellInactive) return;
    const real zero = real(0), dt = (real)pl.dt;
    real req[3];
    bool donor[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int en = ent[k];
        req[k] = qk[k] * dt;
        // interior entry, and this cell is its donor: the from (left) cell when the
        // request is >= 0, the to (right) cell otherwise
        donor[k] = en >= 0 && !(en & 2) && ((req[k] >= zero) != ((en & 1) != 0));
    }
    real sum = zero;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const unsigned k = (order >> (2 * j)) & 3u;
        if (k == 3u) break;
        if (k == 0u ? donor[0] : (k == 1u ? donor[1] : donor[2])) {
            const real r = k == 0u ? req[0] : (k == 1u ? req[1] : req[2]);
            sum = sum + fabs(r);
        }
    }
    const real factor = vmin(real(1), Divisor<real>(vmax(sum, p.tiny)).div(vmax(h, zero) * A));
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (donor[k]) T[(ent[k] >> 2) + eb] = (req[k] >= zero ? real(1) : real(-1)) * (fabs(req[k]) * factor);
}

// Error-free transformation a + b = s + err (TwoSum, Knuth).
template <typename real>
__device__ __forceinline__ real two_sum(real a, real b, real& err) {
    const real s = a + b;
    const real bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
}

// ---------------------------------------------------------------------------
// k_update: one thread per cell.  Gathers its (<= 3) edge records through the
// cell->edge table (left-side entries in ascending edge order, then right-side
// entries), applies the mass/momentum update and Manning friction
// (v1_parity.py:203-209), rain + Horton (:217-242) and the velocity clamp
// (:245-256), and the pending wet-time increment of the previous step
// (:461-463).  COMP carries a per-cell TwoSum compensation of h, hu, hv.
// ---------------------------------------------------------------------------
// Per-cell source parameters of the update: Manning n^2, rain multiplier (with
// the active mask folded in) and the Horton f0, fc, k.  Either one array each,
// or -- when the batch has at most 65,535 distinct (n^2, m, f0, fc, k) tuples,
// as land-cover classes give -- a 16-bit class per cell and a class table
// {n^2, m, f0, fc} + k, which holds the same values in 2 bytes per cell.
template <typename real>
struct CellSources {
    const real *n2, *rainm, *f0, *fc, *hk;
    const unsigned short* cls;
    const typename V4<real>::type* tab;
    const real* tab_hk;
};

// LI: the local-inertial surface (triangular_backend.py:1249-1286) in place of
// the flux divergence and friction: rec holds {Q, Q nx, Q ny} per edge and recR
// the donor-limited transfer T of each interior edge (k_li_flux, k_li_donor).
// SH: the cell->edge table is shared (as in k_li_donor).
template <typename real, bool COMP, bool CLS, bool LI, bool SH, bool GATE>
__global__ void __launch_bounds__(256) k_update(Ctl* __restrict__ ctls, KP<real> p, Seg seg, int parity,
        ForcingTables ft,
        const int* __restrict__ ce0, const int* __restrict__ ce1, const int* __restrict__ ce2,
        const typename V4<real>::type* __restrict__ rec, const real* __restrict__ recR,
        typename V4<real>::type* __restrict__ state, typename V4<real>::type* __restrict__ comp,
        real* __restrict__ wet, const real* __restrict__ area_, const real* __restrict__ lc_,
        CellSources<real> src, double* __restrict__ part) {
    typedef typename V4<real>::type V;
    const int i = xcd_block(blockIdx.x, gridDim.x, p.remap);
    int b, m, c0 = 0, eb = 0;
    if constexpr (SH) {
        b = shared_block(seg, i, m);
        c0 = m * seg.sh_items;
        eb = m * seg.sh_link;
    } else {
        b = i;
        m = seg.blk_member[b];
    }
    // (the member's first block goes on: it carries the clock to the next step's slot)
    if (GATE && !step_runs(ctls + m, parity) && b != seg.mem_blk0[m]) return;
    const int c = b * blockDim.x + threadIdx.x;
    const real zero = real(0), eps = p.eps;
    // every load of the cell (the grid covers the padded cell space exactly;
    // padding and inactive cells carry no table entries) ahead of the control block
    V s = state[c];
    const real wt0 = wet[c];
    const int ent[3] = {ce0[c - c0], ce1[c - c0], ce2[c - c0]};
    // The three gathers are unconditional (a missing entry reads record 0 and
    // is ignored below), so they issue back to back with no branch between
    // them; recR is gathered for left entries too and ignored there.
    V rr[3];
    real rR[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int en = ent[k];
        const int pos = en >= 0 ? (en >> 2) + eb : 0;
        if constexpr (LI) {
            // the local-inertial records are {Q, Q nx, Q ny}, 12 bytes (k_li_flux)
            const Rec3<real> t = reinterpret_cast<const Rec3<real>*>(rec)[pos];
            rr[k].x = t.x; rr[k].y = t.y; rr[k].z = t.z; rr[k].w = real(0);
        } else {
            rr[k] = rec[pos];
        }
        rR[k] = recR[pos];
    }
    const real A = area_[c];
    const real lcc = LI ? lc_[c] : real(0);
    real n2, rm, f0, fc, hk;
    if (CLS) {
        const int k = src.cls[c];
        const V t = src.tab[k];
        n2 = t.x; rm = t.y; f0 = t.z; fc = t.w;
        hk = src.tab_hk[k];
    } else {
        n2 = src.n2[c]; rm = src.rainm[c]; f0 = src.f0[c]; fc = src.fc[c]; hk = src.hk[c];
    }
    V cc;
    if (COMP) cc = comp[c];
    SlotAcc<4> acc4;
    acc4.prefetch(part, b);
    // the member's block range, read before the barrier (no store can precede
    // it, so these are scalar loads)
    const int blk0 = seg.mem_blk0[m], mend = seg.mem_end[m];
    FH_ISSUE_LOADS_OF("s"(state), "s"(wet), "s"(ce0), "s"(ce1), "s"(ce2), "s"(rec), "s"(recR), "s"(area_),
                      "s"(comp), "s"(part));
    // the step plan (dt policy, rain, inflow interpolation) from the start-of-step
    // clock, the same in every block; the member's first block commits it
    Ctl* ctl = ctls + m;
    const Clock ck = ctl->clk[parity];
    const bool eval = ck.active && ck.eval && !ctl->override_dt && !(p.fixed_dt > 0.0);
    const double gmin = eval ? cfl_slots_min(p.cfl_slots, m, parity) : INFINITY;
    const StepPlan<real> pl = plan_step<real>(ctl, ck, p, ft, m, gmin);
    const bool committer = b == blk0 && threadIdx.x == 0;
    if (!pl.active) {
        if (committer) commit_plan<real>(ctl, parity, ck, pl, p, ft, m, gmin);
        return;
    }
    const real dt = (real)pl.dt;
    const real dt_wet = (real)pl.dt_wet;
    const real rate = (real)pl.rate;
    double rain_v = 0.0, inf_v = 0.0, bin = 0.0, bout = 0.0;
    if (c < mend) {
        // deferred wet-time increment of the previous step (h here is its final depth)
        real wt = wt0;
        wt = wt + (s.x > eps ? dt_wet : zero);
        const int e0 = ent[0];
        if (e0 == kCellInactive) {
            if constexpr (!LI) {
                // code-3 boundary edges of the inactive cell (slots 1 and 2): their
                // stage inflow counts as boundary volume, as for an active cell
#pragma unroll
                for (int k = 1; k < 3; ++k) {
                    if (ent[k] < 0 || !(ent[k] & 2)) continue;
                    const double outward = (double)(rr[k].x * dt);
                    bin += fmax(-outward, 0.0);
                    bout += fmax(outward, 0.0);
                }
            }
            s.x = zero; s.y = zero; s.z = zero;
            state[c] = s;
            wet[c] = wt;
            if (COMP) { V z; z.x = zero; z.y = zero; z.z = zero; z.w = zero; comp[c] = z; }
        } else {
            real h, hu, hv;
            real ch = zero, chu = zero, chv = zero;
            if constexpr (LI) {
                // continuity: -T at the from entries, then +T at the to entries
                // (network_kernels.py:70-82), over the interior edges
                real delta = zero;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int en = ent[k];
                    if (en < 0 || (en & 2)) continue;
                    delta = delta + ((en & 1) ? rR[k] : -rR[k]);
                }
                const real vol_after = vmax(vmax(s.x, zero) * A + delta, zero);
                // boundary outflow max(Q, 0) dt at this cell's boundary edges (zero at
                // walls), limited by the volume left (triangular_backend.py:1767-1772)
                real req[3];
                real bsum = zero;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int en = ent[k];
                    req[k] = (en >= 0 && (en & 2)) ? vmax(rr[k].x, zero) * dt : zero;
                    if (en >= 0 && (en & 2)) bsum = bsum + req[k];
                }
                const real bfac = vmin(real(1), Divisor<real>(vmax(bsum, p.tiny)).div(vol_after));
                real out = zero;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int en = ent[k];
                    if (en < 0 || !(en & 2)) continue;
                    const real o = req[k] * bfac;
                    out = out + o;
                    bout += (double)o;
                }
                const real vdelta = delta - out;
                if (COMP) {
                    h = two_sum<real>(s.x, Divisor<real>(A).div(vdelta) + cc.x, ch);
                    if (!(h > zero)) { h = vmax(h, zero); ch = zero; }
                } else {
                    h = vmax(s.x + Divisor<real>(A).div(vdelta), zero);
                }
                // momentum from the edge discharges around the cell
                // (triangular_backend.py:1270-1285): from entries, then to entries
                real Lu = zero, Lv = zero, Ru = zero, Rv = zero;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int en = ent[k];
                    if (en < 0) continue;
                    if (en & 1) { Ru = Ru + rr[k].y; Rv = Rv + rr[k].z; }
                    else { Lu = Lu + rr[k].y; Lv = Lv + rr[k].z; }
                }
                const real perimeter = Divisor<real>(vmax(lcc, p.tiny)).div(real(2) * A);
                const bool w = h > eps;
                const Divisor<real> dp(vmax(perimeter, p.tiny));
                hu = w ? dp.div(Lu + Ru) : zero;
                hv = w ? dp.div(Lv + Rv) : zero;
            } else {
            real Lh = zero, Lu = zero, Lv = zero, Rh = zero, Ru = zero, Rv = zero;
            bool anyR = false;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int en = ent[k];
                if (en < 0) continue;
                const V r = rr[k];
                if (en & 1) {
                    Rh = Rh + r.x; Ru = Ru + r.w; Rv = Rv + rR[k]; anyR = true;
                } else {
                    Lh = Lh + (-r.x); Lu = Lu + (-r.y); Lv = Lv + (-r.z);
                    if (en & 2) {      // boundary edge: outward volume of this step (v1_parity.py:211-213)
                        const double outward = (double)(r.x * dt);
                        bin += fmax(-outward, 0.0);
                        bout += fmax(outward, 0.0);
                    }
                }
            }
            const real dh = anyR ? Lh + Rh : Lh;
            const real dhu = anyR ? Lu + Ru : Lu;
            const real dhv = anyR ? Lv + Rv : Lv;
            const Divisor<real> dA(A);
            if (COMP) {
                h = two_sum<real>(s.x, dA.div(dt * dh) + cc.x, ch);
                hu = two_sum<real>(s.y, dA.div(dt * dhu) + cc.y, chu);
                hv = two_sum<real>(s.z, dA.div(dt * dhv) + cc.z, chv);
                if (!(h > zero)) { h = vmax(h, zero); ch = zero; }
            } else {
                h = vmax(s.x + dA.div(dt * dh), zero);
                hu = s.y + dA.div(dt * dhu);
                hv = s.z + dA.div(dt * dhv);
            }
            // Manning friction (semi-implicit)
            {
                const bool w = h > eps;
                const real hs = vmax(h, eps);
                const Divisor<real> dhs(hs);
                const real u = dhs.div(hu), v = dhs.div(hv);
                const real speed = sqrt(u * u + v * v);
                const real damping = real(1) + Divisor<real>(fric_pow(hs, p.pow43_dln2)).div(dt * p.g * n2 * speed);
                if (w) {
                    const Divisor<real> dd(damping);
                    hu = dd.div(hu); hv = dd.div(hv);
                    if (COMP) { chu = dd.div(chu); chv = dd.div(chv); }
                } else { hu = zero; hv = zero; chu = zero; chv = zero; }
            }
            }
            // rain + Horton infiltration (capacity from the wet time of the previous step)
            const real rain_depth = vmax(rate, zero) * rm * dt;
            real h_before, e1 = zero;
            if (COMP) h_before = two_sum<real>(h, rain_depth, e1); else h_before = h + rain_depth;
            const real capacity = fc + (f0 - fc) * exp(-hk * wt);
            const real infil = vmin(h_before, vmax(capacity, zero) * dt);
            real h2, e2 = zero;
            if (COMP) h2 = two_sum<real>(h_before, -infil, e2); else h2 = h_before - infil;
            const real scale = h_before > eps ? Divisor<real>(vmax(h_before, eps)).div(h2) : zero;
            if (h2 > eps) { hu = hu * scale; hv = hv * scale; if (COMP) { chu = chu * scale; chv = chv * scale; } }
            else { hu = zero; hv = zero; chu = zero; chv = zero; }
            // velocity clamp
            const bool w2 = h2 > eps;
            const real speed2 = w2 ? Divisor<real>(vmax(h2, eps)).div(sqrt(hu * hu + hv * hv)) : zero;
            const real sc = vmin(real(1), Divisor<real>(vmax(speed2, p.tiny)).div(p.max_speed));
            if (w2) { hu = hu * sc; hv = hv * sc; if (COMP) { chu = chu * sc; chv = chv * sc; } }
            else { hu = zero; hv = zero; chu = zero; chv = zero; }
            s.x = h2; s.y = hu; s.z = hv;
            state[c] = s;
            wet[c] = wt;
            if (COMP) {
                V cc;
                cc.x = (h2 > zero) ? (ch + e1) + e2 : zero;
                cc.y = chu; cc.z = chv; cc.w = zero;
                comp[c] = cc;
            }
            rain_v = (double)rain_depth * (double)A;
            inf_v = (double)infil * (double)A;
        }
    }
    __shared__ double lds[16 * 4];
    double v[4] = {rain_v, inf_v, bin, bout};
    block_sum<4>(v, lds);
    acc4.commit(part, b, v);
    if (committer) commit_plan<real>(ctl, parity, ck, pl, p, ft, m, gmin);
}

// Applies the pending wet-time increment of each member's last executed step
// (the clock of parity `parity` is the next step's start).
template <typename real>
__global__ void k_finalize_wet(const Ctl* __restrict__ ctls, KP<real> p, Seg seg, int parity,
                               const typename V4<real>::type* __restrict__ state, real* __restrict__ wet) {
    const int b = blockIdx.x;
    const int m = seg.blk_member[b];
    const Clock& ck = ctls[m].clk[parity];
    if (!ck.pending_wet) return;
    const int c = b * blockDim.x + threadIdx.x;
    if (c >= seg.mem_end[m]) return;
    const real dt = (real)ck.dt_prev;
    if (state[c].x > p.eps) wet[c] = wet[c] + dt;
}

__global__ void k_clear_pending(Ctl* ctls, int members, int parity) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < members) ctls[m].clk[parity].pending_wet = 0;
}

// Starts an advance: target, event end and step budget; the clock of the next
// step (parity `parity`) starts a new budget and decides whether it executes.
__global__ void k_set_target(Ctl* ctls, int members, int parity, double target, double final_t, int64_t max_steps) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members) return;
    Ctl* ctl = ctls + m;
    ctl->target = target;
    ctl->final_t = final_t;
    ctl->max_steps = max_steps;
    Clock& ck = ctl->clk[parity];
    ck.interval_step = 0;
    ck.active = (ck.t < target) && (0 < max_steps);
    if (m == 0) ctl->nonfinite = 0;
}

__global__ void k_set_override(Ctl* ctls, int members, int on, double dt, double rain) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members) return;
    Ctl* ctl = ctls + m;
    ctl->override_dt = on;
    ctl->override_dt_value = dt;
    ctl->override_rain = rain;
}

// Ledger accumulation slots of one reducing kernel: the member's block range
// in that kernel's index space, N terms per slot, and the MassLedger terms.
struct FoldGroup {
    double* acc;
    const int* blk0;    // [M] first block of the member
    const int* nblk;    // [M] number of blocks
    int n;              // terms per slot (1..4)
    int terms[4];
};

struct FoldTables {
    FoldGroup g[6];
    int ngroups;
};

// Sum of one member's slots of one group (N terms each) into v (thread 0), in a
// fixed order, then the slots cleared.  The slot loads carry no store between
// them, so each thread's loads of an unrolled stretch issue together.
constexpr int kFoldThreads = 1024;
template <int N>
__device__ __forceinline__ void fold_group(double* __restrict__ acc, int b0, int nb, double (&v)[4], double* lds) {
    double s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = 0.0;
    const double* src = acc + (size_t)b0 * N;
#pragma unroll 4
    for (int i = threadIdx.x; i < nb; i += kFoldThreads) {
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] += src[(size_t)i * N + k];
    }
    block_sum<N>(s, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = s[k];
    }
    for (int i = threadIdx.x; i < nb * N; i += kFoldThreads) acc[(size_t)b0 * N + i] = 0.0;
}

// k_fold: block (member m, group g) sums the member's accumulation slots of
// group g (fixed order), writes the sums as the member's diagnostics of the
// group's ledger terms, adds them to the member's FP64 ledger and clears the
// slots.  The groups hold disjoint terms; the boundary net term is the first
// group's in - out.  Blocks from f.ngroups * members on check the batch's depths
// and node volumes for non-finite values (parity_runner.py:319-322 checks),
// one item per thread; every thread that finds one stores 1 in ctls[0].nonfinite.
template <typename real>
__global__ void __launch_bounds__(kFoldThreads) k_fold(Ctl* __restrict__ ctls, FoldTables f, int members,
        int n_cells, const typename V4<real>::type* __restrict__ state, int n_nodes, const real* __restrict__ vol) {
    __shared__ double lds[16 * 4];
    const int blk = blockIdx.x;
    if (blk >= f.ngroups * members) {
        const int i = (blk - f.ngroups * members) * kFoldThreads + threadIdx.x;
        if (i >= n_cells + n_nodes) return;
        const real v = i < n_cells ? state[i].x : vol[i - n_cells];
        if (!isfinite(v)) ctls[0].nonfinite = 1;
        return;
    }
    const int m = blk / f.ngroups, gi = blk % f.ngroups;
    Ctl* ctl = ctls + m;
    const FoldGroup& g = f.g[gi];
    const int b0 = g.blk0[m], nb = g.nblk[m];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    switch (g.n) {
        case 1: fold_group<1>(g.acc, b0, nb, v, lds); break;
        case 2: fold_group<2>(g.acc, b0, nb, v, lds); break;
        case 3: fold_group<3>(g.acc, b0, nb, v, lds); break;
        default: fold_group<4>(g.acc, b0, nb, v, lds); break;
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < g.n; ++k) {
            ctl->diag[g.terms[k]] = v[k];
            ctl->ledger[g.terms[k]] += v[k];
        }
        if (gi == 0) {
            const double bnet = v[2] - v[3];   // the update group's terms: rain, infiltration, in, out
            ctl->diag[L_BNET] = bnet;
            ctl->ledger[L_BNET] += bnet;
        }
    }
}

}  // namespace fh
