/*
 * CPU oracle, C restatement: one V1-parity solver step on all host cores.
 *
 * TEST INFRASTRUCTURE ONLY (with oracle/v1_parity_np.py, which it restates
 * operation by operation).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, as the checker or the timed CPU baseline.
 *
 * Each function restates the NumPy oracle's function of the same name, which in
 * turn restates the reference (paths relative to
 * HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/):
 *   hll, hydrostatic_flux_pair   hybrid_flood/jax_solver/v1_parity.py:110-150
 *   stable_timestep              v1_parity.py:153-158
 *   surface_flux_step            v1_parity.py:161-214 (+ the V1 stage ghost,
 *                                gurugram_flood/kernels/triangular_backend.py:1294-1327)
 *   rainfall_horton              v1_parity.py:217-242
 *   clamp_velocity               v1_parity.py:245-256
 *   inlet_exchange               v1_parity.py:259-302
 *   pipe_links_and_outfalls      v1_parity.py:305-368
 *   return_node_overflow         v1_parity.py:371-404
 *   external_inflow, recharge    v1_parity.py:407-427
 *   oracle_step                  v1_parity.py:430-469 (V1 process order)
 *   li_stable_timestep,          the local-inertial surface mode (oracle/local_inertial_np.py):
 *   li_surface_step              gurugram_flood/kernels/triangular_backend.py:1181-1291,
 *                                local_inertial.py:20-48, network_kernels.py:45-82,
 *                                triangular_backend.py:1767-1772, equations.py:181-207
 *
 * Arithmetic: every operation in REAL (float or double; the file is compiled
 * once for each), rounded per operation (-ffp-contract=off, no fast-math), in
 * the NumPy oracle's operation order.  max/min propagate NaN like
 * numpy.maximum/minimum.  Every scatter-add of the oracle (np.add.at,
 * sequential in index order) is either the same sequential loop or, for the
 * surface, a per-cell gather over the cell's edges in ascending edge order
 * (left-side edges, then right-side edges), which sums the same operands in the
 * same order.  Ledger terms are FP64 sums in a fixed chunk order, so results do
 * not depend on the thread count.  Differences from the NumPy oracle come only
 * from libm's exp/pow/acos/sin against NumPy's own float32 routines.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif
#ifndef SFX
#define SFX _f32
#endif
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define API(name) CAT(name, SFX)

#define GRAVITY 9.80665
#define MIN_DEPTH 1.0e-5
#define NCHUNK 256 /* fixed summation chunks of the FP64 ledger sums */
#define PI_D 3.14159265358979323846 /* numpy.pi */

typedef REAL R;

static inline R vmax(R a, R b) { return (a >= b || a != a) ? a : b; }
static inline R vmin(R a, R b) { return (a <= b || a != a) ? a : b; }
static inline R sgn(R x) { return x > (R)0 ? (R)1 : (x < (R)0 ? (R)-1 : x); }

#if defined(REAL_IS_DOUBLE)
#define SQRT sqrt
#define EXP exp
#define POW pow
#define ACOS acos
#define SIN sin
#define FABS fabs
#else
#define SQRT sqrtf
#define EXP expf
#define POW powf
#define ACOS acosf
#define SIN sinf
#define FABS fabsf
#endif

/* The topology, forcing fields and gather tables, all host arrays owned by the
 * caller (ctypes).  Field meanings are the reference NamedTuples'. */
typedef struct {
    int64_t C, E, I, N, L, S, Rr, X;
    /* surface */
    const R *bed, *manning, *area, *lc;
    const int32_t *edge_left, *edge_right, *code;
    const R *nx, *ny, *len;
    const uint8_t* active;
    /* cell -> edge gathers: left-side edges of a cell, ascending; right-side
     * edges of a cell over active interior edges, ascending */
    const int64_t *lptr, *rptr;
    const int32_t *lidx, *ridx;
    /* stage depth per edge (code-3 edges), or NULL: no stage values */
    const R* stage;
    /* coupling */
    const int32_t *inlet_cell, *inlet_node, *node_cell;
    const R *inv, *As, *dmax, *capq, *surq;
    const int32_t *scell, *snode;
    const R* sw;
    int32_t snorm;
    /* pipes */
    const int32_t *lfrom, *lto, *geom;
    const R *llen, *lman, *lcap, *linv, *D, *minor, *rw, *rh, *oarea, *ocoef, *otw, *ocap;
    /* recharge */
    const int32_t* rcell;
    const R *rS, *rcc, *rrc, *rstop;
    /* forcing fields */
    const R *rainm, *f0, *fc, *hk;
    const int32_t* rcv;
    /* surface operator: 0 the full shallow-water step, 1 the local-inertial step */
    int32_t li;
} Topo;

/* Scratch buffers (caller-allocated, reused every step). */
typedef struct {
    R *fl0, *fl1, *fl2, *fr0, *fr1, *fr2; /* per edge, times length */
    double *bpart;                        /* per edge outward volume (boundary edges) */
    R *cell_tmp, *node_tmp, *node_tmp2, *head, *inlet_a, *inlet_b, *link_a, *link_b, *entry_a;
    uint8_t* link_pos;
    double *sum_a, *sum_b;                /* max(E, C) each: operands of the FP64 ledger sums */
    R *cell_b, *cell_c;                   /* per cell (local-inertial step) */
} Work;

static void chunk_sum(const double* v, int64_t n, double* out) {
    double part[NCHUNK];
    const int64_t step = (n + NCHUNK - 1) / NCHUNK;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < NCHUNK; ++k) {
        double s = 0.0;
        const int64_t a = k * step, b = a + step < n ? a + step : n;
        for (int64_t i = a; i < b; ++i) s += v[i];
        part[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < NCHUNK; ++k) s += part[k];
    *out = s;
}

/* v1_parity.py:110-132 */
static void hll(R h_l, R hu_l, R hv_l, R h_r, R hu_r, R hv_r, R nx, R ny, R out[3]) {
    const R g = (R)GRAVITY, dry = (R)MIN_DEPTH, zero = (R)0;
    h_l = vmax(h_l, zero);
    h_r = vmax(h_r, zero);
    const int wl = h_l > dry, wr = h_r > dry;
    const R u_l = wl ? hu_l / vmax(h_l, dry) : zero;
    const R v_l = wl ? hv_l / vmax(h_l, dry) : zero;
    const R u_r = wr ? hu_r / vmax(h_r, dry) : zero;
    const R v_r = wr ? hv_r / vmax(h_r, dry) : zero;
    const R un_l = u_l * nx + v_l * ny, un_r = u_r * nx + v_r * ny;
    const R c_l = SQRT(g * h_l), c_r = SQRT(g * h_r);
    const R half_g = (R)(0.5 * GRAVITY);
    const R p_l = half_g * (h_l * h_l), p_r = half_g * (h_r * h_r);
    const R fl[3] = {h_l * un_l, hu_l * un_l + p_l * nx, hv_l * un_l + p_l * ny};
    const R fr[3] = {h_r * un_r, hu_r * un_r + p_r * nx, hv_r * un_r + p_r * ny};
    const R ql[3] = {h_l, hu_l, hv_l}, qr[3] = {h_r, hu_r, hv_r};
    const R s_l = vmin(un_l - c_l, un_r - c_r);
    const R s_r = vmax(un_l + c_l, un_r + c_r);
    const R denom = FABS(s_r - s_l) > (R)1.0e-12 ? s_r - s_l : (R)1;
    for (int k = 0; k < 3; ++k) {
        R f;
        if (s_l >= zero) f = fl[k];
        else if (s_r <= zero) f = fr[k];
        else f = (s_r * fl[k] - s_l * fr[k] + s_l * s_r * (qr[k] - ql[k])) / denom;
        out[k] = (wl || wr) ? f : zero;
    }
}

/* v1_parity.py:135-150 */
static void hydrostatic_flux_pair(R h_l, R hu_l, R hv_l, R z_l, R h_r, R hu_r, R hv_r, R z_r, R nx, R ny,
                                  R fl[3], R fr[3]) {
    const R zero = (R)0, eps = (R)MIN_DEPTH;
    h_l = vmax(h_l, zero);
    h_r = vmax(h_r, zero);
    const R z_star = vmax(z_l, z_r);
    const R hs_l = vmax(h_l + z_l - z_star, zero);
    const R hs_r = vmax(h_r + z_r - z_star, zero);
    const int wl = h_l > eps, wr = h_r > eps;
    const R u_l = wl ? hu_l / vmax(h_l, eps) : zero;
    const R v_l = wl ? hv_l / vmax(h_l, eps) : zero;
    const R u_r = wr ? hu_r / vmax(h_r, eps) : zero;
    const R v_r = wr ? hv_r / vmax(h_r, eps) : zero;
    R f[3];
    hll(hs_l, u_l * hs_l, v_l * hs_l, hs_r, u_r * hs_r, v_r * hs_r, nx, ny, f);
    const R half_g = (R)(0.5 * GRAVITY);
    const R pc_l = half_g * (h_l * h_l - hs_l * hs_l);
    const R pc_r = half_g * (h_r * h_r - hs_r * hs_r);
    fl[0] = f[0] + zero; fl[1] = f[1] + pc_l * nx; fl[2] = f[2] + pc_l * ny;
    fr[0] = f[0] + zero; fr[1] = f[1] + pc_r * nx; fr[2] = f[2] + pc_r * ny;
}

static R li_stable_timestep(const Topo* t, const R* h, double cfl);

/* v1_parity.py:153-158: min over cells of cfl Lc / (|u| + sqrt(g h)), non-finite skipped
 * (the local-inertial mode: its gravity-wave timestep) */
R API(oracle_stable_timestep)(const Topo* t, const R* h, const R* hu, const R* hv, double cfl) {
    if (t->li) return li_stable_timestep(t, h, cfl);
    const R eps = (R)MIN_DEPTH, c = (R)cfl;
    R best = (R)INFINITY;
#pragma omp parallel
    {
        R mn = (R)INFINITY;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < t->C; ++i) {
            const R hs = vmax(h[i], eps);
            const R u = hu[i] / hs, v = hv[i] / hs;
            R speed = SQRT(u * u + v * v);
            speed = speed + SQRT((R)GRAVITY * hs);
            R local = c * t->lc[i] / vmax(speed, (R)1.0e-12);
            if (!isfinite(local)) local = (R)INFINITY;
            if (local < mn) mn = local;
        }
#pragma omp critical
        if (mn < best) best = mn;
    }
    return best;
}

/* equations.py:181-207: min over cells of cfl Lc / max(sqrt(g max(h, min depth)), 1e-12) */
static R li_stable_timestep(const Topo* t, const R* h, double cfl) {
    const R c = (R)cfl;
    R best = (R)INFINITY;
#pragma omp parallel
    {
        R mn = (R)INFINITY;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < t->C; ++i) {
            const R wave = SQRT((R)GRAVITY * vmax(h[i], (R)MIN_DEPTH));
            R local = c * t->lc[i] / vmax(wave, (R)1.0e-12);
            if (!isfinite(local)) local = (R)INFINITY;
            if (local < mn) mn = local;
        }
#pragma omp critical
        if (mn < best) best = mn;
    }
    return best;
}

/* local_inertial_np.surface_step (triangular_backend.py:1181-1291).  Every
 * scatter-add of the NumPy form is a per-cell gather in ascending edge order:
 * the donor sums merge the cell's left-side and right-side edge lists. */
static void li_surface_step(const Topo* t, Work* w, R* h, R* hu, R* hv, R dt, double* diag) {
    const R zero = (R)0, half = (R)0.5, g = (R)GRAVITY, eps = (R)MIN_DEPTH;
    const R gdt = (R)(GRAVITY * (double)dt);
    R *q = w->fl0, *T = w->fl1, *outq = w->fl2, *qx = w->fr0, *qy = w->fr1;
    R *factor = w->cell_tmp, *idelta = w->cell_b, *cfac = w->cell_c;
    /* edge discharge (local_inertial.py:20-48) and the requested transfers */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < t->E; ++e) {
        const int32_t left = t->edge_left[e], right = t->edge_right[e];
        const int interior = right >= 0;
        const int32_t sr = interior ? right : left;
        const int la = t->active[left] != 0, ra = t->active[sr] != 0;
        const int ai = interior && la && ra;
        const int32_t code = (interior && la && !ra) ? 0 : t->code[e];
        const R h_l = la ? h[left] : zero, hu_l = la ? hu[left] : zero, hv_l = la ? hv[left] : zero;
        const R h_r = ra ? h[sr] : zero, hu_r = ra ? hu[sr] : zero, hv_r = ra ? hv[sr] : zero;
        const R bed_l = t->bed[left], bed_r = t->bed[sr];
        const R eta_l = bed_l + h_l;
        const R eta_r = ai ? bed_r + h_r : bed_l;
        const R bed_star = ai ? vmax(bed_l, bed_r) : bed_l;
        const R h_flow = vmax(vmax(eta_l, eta_r) - bed_star, zero);
        const R L = t->len[e], nx = t->nx[e], ny = t->ny[e];
        const R distance = vmax(half * (t->lc[left] + t->lc[sr]), (R)1.0e-6);
        const R slope = (eta_r - eta_l) / distance;
        const R nm_l = hu_l * nx + hv_l * ny, nm_r = hu_r * nx + hv_r * ny;
        const R q_old = (ai ? half * (nm_l + nm_r) : nm_l) * L;
        const R n_face = ai ? half * (t->manning[left] + t->manning[sr]) : t->manning[left];
        const int bedge = !interior && la && (code == 4 || code == 1);
        const R area = vmax(h_flow * L, (R)1.0e-8);
        const R radius = vmax(h_flow, (R)1.0e-6);
        const R numerator = q_old - g * area * dt * slope;
        const R denominator = (R)1 + (gdt * (n_face * n_face) * FABS(q_old) / (POW(radius, (R)(4.0 / 3.0)) * area));
        R qq = numerator / denominator;
        qq = (ai || bedge) ? qq : zero;
        if (bedge) qq = vmax(qq, zero);
        q[e] = qq;
        T[e] = ai ? qq * dt : zero;
        outq[e] = bedge ? vmax(qq, zero) * dt : zero;
        qx[e] = qq * nx;
        qy[e] = qq * ny;
    }
    /* donor factors (network_kernels.py:45-67): the cell's |request| over the
     * edges it donates to, in ascending edge order */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) {
        R sum = zero;
        int64_t a = t->lptr[c], b = t->rptr[c];
        const int64_t a1 = t->lptr[c + 1], b1 = t->rptr[c + 1];
        while (a < a1 || b < b1) {
            const int takel = b >= b1 || (a < a1 && t->lidx[a] < t->ridx[b]);
            const int32_t e = takel ? t->lidx[a++] : t->ridx[b++];
            const R r = T[e];
            if (takel ? r >= zero : r < zero) sum = sum + FABS(r);
        }
        const R volume = vmax(h[c], zero) * t->area[c];
        factor[c] = vmin((R)1, volume / vmax(sum, (R)1.0e-12));
    }
    /* limited signed transfers */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < t->E; ++e) {
        const R r = T[e];
        const int32_t left = t->edge_left[e], right = t->edge_right[e];
        const int32_t donor = r >= zero ? left : (right >= 0 ? right : left);
        T[e] = (r >= zero ? (R)1 : (R)-1) * (FABS(r) * factor[donor]);
    }
    /* continuity (network_kernels.py:70-82: all 'from' entries, then all 'to'
     * entries, one running sum) and the capture factor of the boundary outflow
     * (triangular_backend.py:1767-1772) */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) {
        R acc = zero, by = zero;
        for (int64_t k = t->lptr[c]; k < t->lptr[c + 1]; ++k) acc = acc + (-T[t->lidx[k]]);
        for (int64_t k = t->rptr[c]; k < t->rptr[c + 1]; ++k) acc = acc + T[t->ridx[k]];
        for (int64_t k = t->lptr[c]; k < t->lptr[c + 1]; ++k) by = by + vmax(outq[t->lidx[k]], zero);
        idelta[c] = acc;
        const R volume_after = vmax(vmax(h[c], zero) * t->area[c] + acc, zero);
        cfac[c] = vmin((R)1, volume_after / vmax(by, (R)1.0e-12));
    }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < t->E; ++e) {
        outq[e] = vmax(outq[e], zero) * cfac[t->edge_left[e]];
        w->bpart[e] = (double)outq[e];
    }
    /* depth and the momentum rebuilt from the edge discharges */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) {
        R out = zero, sx = zero, sy = zero, rx = zero, ry = zero;
        for (int64_t k = t->lptr[c]; k < t->lptr[c + 1]; ++k) {
            const int32_t e = t->lidx[k];
            out = out + outq[e];
            sx = sx + qx[e];
            sy = sy + qy[e];
        }
        for (int64_t k = t->rptr[c]; k < t->rptr[c + 1]; ++k) {
            const int32_t e = t->ridx[k];
            rx = rx + qx[e];
            ry = ry + qy[e];
        }
        const R A = t->area[c];
        const R delta = idelta[c] - out;
        const R hh = t->active[c] ? vmax(h[c] + delta / A, zero) : zero;
        const R perimeter = (R)2 * A / vmax(t->lc[c], (R)1.0e-12);
        const int wet = hh > eps;
        h[c] = hh;
        hu[c] = wet ? (sx + rx) / vmax(perimeter, (R)1.0e-12) : zero;
        hv[c] = wet ? (sy + ry) / vmax(perimeter, (R)1.0e-12) : zero;
    }
    double bout = 0.0;
    chunk_sum(w->bpart, t->E, &bout);
    diag[0] = 0.0 - bout;
    diag[1] = 0.0;
    diag[2] = bout;
}

/* v1_parity.py:161-214 (diag: boundary net, in, out) */
static void surface_flux_step(const Topo* t, Work* w, R* h, R* hu, R* hv, R dt, double* diag) {
    const R zero = (R)0, g = (R)GRAVITY, eps = (R)MIN_DEPTH;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < t->E; ++e) {
        const int32_t left = t->edge_left[e], right = t->edge_right[e];
        const int interior = right >= 0;
        const int32_t safe_right = interior ? right : left;
        const int la = t->active[left] != 0, ra = t->active[safe_right] != 0;
        const int ai = interior && la && ra;
        const int inactive_right = interior && la && !ra;
        const int32_t code = inactive_right ? 0 : t->code[e];
        const R h_l = la ? h[left] : zero, hu_l = la ? hu[left] : zero, hv_l = la ? hv[left] : zero;
        const R z_l = t->bed[left];
        const R h_rb = ra ? h[safe_right] : zero, hu_rb = ra ? hu[safe_right] : zero;
        const R hv_rb = ra ? hv[safe_right] : zero, z_rb = t->bed[safe_right];
        const R nx = t->nx[e], ny = t->ny[e];
        const R nm = hu_l * nx + hv_l * ny;
        const R hu_reflect = hu_l - (R)2 * nm * nx;
        const R hv_reflect = hv_l - (R)2 * nm * ny;
        const int reflective = code == 0 || code == 2;
        R h_g = h_l, hu_g = reflective ? hu_reflect : hu_l, hv_g = reflective ? hv_reflect : hv_l;
        if (t->stage && code == 3) {
            h_g = vmax(t->stage[e] - z_l, zero);
            hu_g = zero;
            hv_g = zero;
        }
        const R h_r = ai ? h_rb : h_g, hu_r = ai ? hu_rb : hu_g, hv_r = ai ? hv_rb : hv_g;
        const R z_r = ai ? z_rb : z_l;
        R fl[3], fr[3];
        hydrostatic_flux_pair(h_l, hu_l, hv_l, z_l, h_r, hu_r, hv_r, z_r, nx, ny, fl, fr);
        if (!interior && code == 4 && fl[0] < zero) fl[0] = fl[1] = fl[2] = zero;
        const R L = t->len[e];
        w->fl0[e] = -fl[0] * L; w->fl1[e] = -fl[1] * L; w->fl2[e] = -fl[2] * L;
        w->fr0[e] = ai ? fr[0] * L : zero; w->fr1[e] = ai ? fr[1] * L : zero; w->fr2[e] = ai ? fr[2] * L : zero;
        w->bpart[e] = ai ? 0.0 : (double)(fl[0] * L * dt);
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) {
        R dl[3] = {zero, zero, zero}, dr[3] = {zero, zero, zero};
        for (int64_t k = t->lptr[c]; k < t->lptr[c + 1]; ++k) {
            const int32_t e = t->lidx[k];
            dl[0] = dl[0] + w->fl0[e]; dl[1] = dl[1] + w->fl1[e]; dl[2] = dl[2] + w->fl2[e];
        }
        for (int64_t k = t->rptr[c]; k < t->rptr[c + 1]; ++k) {
            const int32_t e = t->ridx[k];
            dr[0] = dr[0] + w->fr0[e]; dr[1] = dr[1] + w->fr1[e]; dr[2] = dr[2] + w->fr2[e];
        }
        const R d0 = dl[0] + dr[0], d1 = dl[1] + dr[1], d2 = dl[2] + dr[2];
        const int act = t->active[c] != 0;
        const R A = t->area[c];
        R hh = act ? vmax(h[c] + dt * d0 / A, zero) : zero;
        R uu = act ? hu[c] + dt * d1 / A : zero;
        R vv = act ? hv[c] + dt * d2 / A : zero;
        const int wet = hh > eps;
        const R hs = vmax(hh, eps);
        const R u = uu / hs, v = vv / hs;
        const R speed = SQRT(u * u + v * v);
        const R n2 = t->manning[c] * t->manning[c];
        const R damping = (R)1 + dt * g * n2 * speed / POW(hs, (R)(4.0 / 3.0));
        h[c] = hh;
        hu[c] = wet ? uu / damping : zero;
        hv[c] = wet ? vv / damping : zero;
    }
    /* boundary in / out: sums of max(-outward, 0) and max(outward, 0) */
    double bin = 0.0, bout = 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < t->E; ++e) {
        w->sum_a[e] = fmax(-w->bpart[e], 0.0);
        w->sum_b[e] = fmax(w->bpart[e], 0.0);
    }
    chunk_sum(w->sum_a, t->E, &bin);
    chunk_sum(w->sum_b, t->E, &bout);
    diag[0] = bin - bout;
    diag[1] = bin;
    diag[2] = bout;
}

/* v1_parity.py:217-256: rain, Horton, momentum rescale, then the velocity clamp */
static void sources_and_clamp(const Topo* t, Work* w, R* h, R* hu, R* hv, const R* wet_time, R dt, double rate,
                              double max_speed, double* rain_out, double* inf_out) {
    const R zero = (R)0, eps = (R)MIN_DEPTH;
    const R rr = vmax((R)rate, zero);
    double *vr = w->sum_a, *vi = w->sum_b;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) {
        const R act = t->active[c] ? (R)1 : zero;
        const R rain_rate = rr * t->rainm[c];
        const R rain_depth = rain_rate * dt * act;
        const R h_before = h[c] + rain_depth;
        const R capacity = t->fc[c] + (t->f0[c] - t->fc[c]) * EXP(-t->hk[c] * wet_time[c]);
        const R infiltration = vmin(h_before, vmax(capacity, zero) * dt * act);
        const R hh = h_before - infiltration;
        const R scale = h_before > eps ? hh / vmax(h_before, eps) : zero;
        R uu = hh > eps ? hu[c] * scale : zero;
        R vv = hh > eps ? hv[c] * scale : zero;
        /* clamp_velocity */
        const int wet = hh > eps;
        const R speed = wet ? SQRT(uu * uu + vv * vv) / vmax(hh, eps) : zero;
        const R sc = vmin((R)1, (R)max_speed / vmax(speed, (R)1.0e-12));
        h[c] = hh;
        hu[c] = wet ? uu * sc : zero;
        hv[c] = wet ? vv * sc : zero;
        vr[c] = (double)rain_depth * (double)t->area[c];
        vi[c] = (double)infiltration * (double)t->area[c];
    }
    chunk_sum(vr, t->C, rain_out);
    chunk_sum(vi, t->C, inf_out);
}

/* v1_parity.py:407-412 */
static double external_inflow(const Topo* t, Work* w, R* h, const R* q, R dt) {
    if (t->X == 0) return 0.0;
    const R zero = (R)0;
    double total = 0.0;
    memset(w->cell_tmp, 0, sizeof(R) * (size_t)t->C);
    for (int64_t j = 0; j < t->X; ++j) {
        const R added = vmax(q[j], zero) * dt;
        const int32_t c = t->rcv[j];
        w->cell_tmp[c] = w->cell_tmp[c] + added / t->area[c];
        total += (double)added;
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) h[c] = h[c] + w->cell_tmp[c];
    return total;
}

static inline R node_head(const Topo* t, const R* V, int64_t n) {
    return t->inv[n] + vmax(V[n], (R)0) / vmax(t->As[n], (R)1.0e-12);
}

/* v1_parity.py:259-302 */
static void inlet_exchange(const Topo* t, Work* w, R* h, R* V, R dt, double capture_q, double* cap_out,
                           double* sur_out) {
    const R zero = (R)0;
    const int64_t I = t->I, N = t->N, C = t->C;
    R *req_c = w->inlet_a, *req_s = w->inlet_b;
    for (int64_t n = 0; n < N; ++n) w->head[n] = node_head(t, V, n);
    for (int64_t i = 0; i < I; ++i) {
        const int32_t c = t->inlet_cell[i], n = t->inlet_node[i];
        const R sh = t->bed[c] + h[c];
        const R dh = sh - w->head[n];
        const R q = (R)capture_q * SQRT((R)(2.0 * GRAVITY) * FABS(dh)) * sgn(dh);
        req_c[i] = vmin(vmax(q, zero) * dt, vmax(t->capq[i], zero) * dt);
        req_s[i] = vmin(vmax(-q, zero) * dt, vmax(t->surq[i], zero) * dt);
    }
    /* _limit_by_group over cells (available h A) and nodes (available V) */
    R* tot_c = w->cell_tmp;
    memset(tot_c, 0, sizeof(R) * (size_t)C);
    memset(w->node_tmp, 0, sizeof(R) * (size_t)N);
    for (int64_t i = 0; i < I; ++i) {
        req_c[i] = vmax(req_c[i], zero);
        req_s[i] = vmax(req_s[i], zero);
        const int32_t c = t->inlet_cell[i], n = t->inlet_node[i];
        tot_c[c] = tot_c[c] + req_c[i];
        w->node_tmp[n] = w->node_tmp[n] + req_s[i];
    }
    double cap_sum = 0.0, sur_sum = 0.0;
    for (int64_t i = 0; i < I; ++i) {
        const int32_t c = t->inlet_cell[i], n = t->inlet_node[i];
        const R fc = vmin((R)1, (h[c] * t->area[c]) / vmax(tot_c[c], (R)1.0e-12));
        const R fn = vmin((R)1, V[n] / vmax(w->node_tmp[n], (R)1.0e-12));
        req_c[i] = req_c[i] * fc;   /* capture */
        req_s[i] = req_s[i] * fn;   /* surcharge */
        cap_sum += (double)req_c[i];
        sur_sum += (double)req_s[i];
    }
    memset(tot_c, 0, sizeof(R) * (size_t)C);
    memset(w->node_tmp, 0, sizeof(R) * (size_t)N);
    for (int64_t i = 0; i < I; ++i) {
        const int32_t c = t->inlet_cell[i], n = t->inlet_node[i];
        tot_c[c] = tot_c[c] + (req_s[i] - req_c[i]) / t->area[c];
        w->node_tmp[n] = w->node_tmp[n] + (req_c[i] - req_s[i]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < C; ++c) h[c] = vmax(h[c] + tot_c[c], zero);
    for (int64_t n = 0; n < N; ++n) V[n] = vmax(V[n] + w->node_tmp[n], zero);
    *cap_out = cap_sum;
    *sur_out = sur_sum;
}

static void circular_theta(R depth, R diameter, R* radius, R* theta) {
    *radius = (R)0.5 * diameter;
    const R open_depth = vmin(vmax(depth, (R)0), diameter);
    *theta = (R)2 * ACOS(vmin(vmax((*radius - open_depth) / *radius, (R)-1), (R)1));
}

/* v1_parity.py:305-312 */
static R circular_area(R depth, R diameter) {
    diameter = vmax(diameter, (R)1.0e-12);
    R radius, theta;
    circular_theta(depth, diameter, &radius, &theta);
    const R open_area = (R)0.5 * (radius * radius) * (theta - SIN(theta));
    const R full = (R)(0.25 * PI_D) * (diameter * diameter);
    return depth >= diameter ? full : open_area;
}

/* v1_parity.py:315-321 */
static R circular_radius(R depth, R diameter) {
    diameter = vmax(diameter, (R)1.0e-12);
    R radius, theta;
    circular_theta(depth, diameter, &radius, &theta);
    const R perimeter = depth >= diameter ? (R)PI_D * diameter : radius * theta;
    return circular_area(depth, diameter) / vmax(perimeter, (R)1.0e-12);
}

/* v1_parity.py:324-368; *transfer_abs gets sum |transfer| (FP64) */
static double pipe_links_and_outfalls(const Topo* t, Work* w, R* V, R* Q, R dt, double slot_wave_speed,
                                      double* transfer_abs) {
    const R zero = (R)0, g = (R)GRAVITY;
    const int64_t N = t->N, L = t->L;
    *transfer_abs = 0.0;
    if (L == 0 && N == 0) return 0.0;
    if (L > 0) {
        for (int64_t n = 0; n < N; ++n) w->head[n] = node_head(t, V, n);
        R *requested_abs = w->link_a, *factor_T = w->link_b;
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < L; ++l) {
            const int32_t s = t->lfrom[l], d = t->lto[l];
            const R D = t->D[l];
            const R depth = vmax((R)0.5 * (w->head[s] + w->head[d]) - t->linv[l], zero);
            const R Dm = vmax(D, (R)1.0e-12);
            const R slot_width = g * ((R)(0.25 * PI_D) * (Dm * Dm)) / (R)(slot_wave_speed * slot_wave_speed);
            R flow_area = circular_area(depth, D) + slot_width * vmax(depth - D, zero);
            const R circ_radius = circular_radius(depth, D);
            const R open_depth = vmin(vmax(depth, zero), vmax(t->rh[l], (R)1.0e-6));
            const R open_area = vmax(t->rw[l], (R)1.0e-6) * open_depth;
            const R open_radius = open_area / vmax(t->rw[l] + (R)2 * open_depth, (R)1.0e-12);
            const int rect = t->geom[l] == 2;
            flow_area = rect ? open_area : flow_area;
            R radius = rect ? open_radius : circ_radius;
            const R length = vmax(t->llen[l], (R)1.0e-6);
            const R area = vmax(flow_area, (R)1.0e-12);
            radius = vmax(radius, (R)1.0e-12);
            const R slope = (w->head[s] - w->head[d]) / length;
            const R nman = vmax(t->lman[l], zero);
            const R absq = FABS(Q[l]);
            const R friction = g * (nman * nman) * absq / (POW(radius, (R)(4.0 / 3.0)) * area);
            const R minor = vmax(t->minor[l], zero) * absq / ((R)2 * length * area);
            R cand = (Q[l] + g * area * dt * slope) / ((R)1 + dt * (friction + minor));
            const R cap = vmax(t->lcap[l], zero);
            cand = vmin(vmax(cand, -cap), cap);
            const R req = cand * dt;
            w->link_pos[l] = req >= zero;
            requested_abs[l] = FABS(req);
        }
        R* by_donor = w->node_tmp;
        memset(by_donor, 0, sizeof(R) * (size_t)N);
        for (int64_t l = 0; l < L; ++l) {
            const int32_t donor = w->link_pos[l] ? t->lfrom[l] : t->lto[l];
            by_donor[donor] = by_donor[donor] + requested_abs[l];
        }
        R* fac = w->node_tmp2;
        for (int64_t n = 0; n < N; ++n) fac[n] = vmin((R)1, V[n] / vmax(by_donor[n], (R)1.0e-12));
        double tabs = 0.0;
        for (int64_t l = 0; l < L; ++l) {
            const int32_t donor = w->link_pos[l] ? t->lfrom[l] : t->lto[l];
            factor_T[l] = (w->link_pos[l] ? (R)1 : (R)-1) * requested_abs[l] * fac[donor];
            tabs += (double)FABS(factor_T[l]);
        }
        /* node_delta: -T over the 'from' ends in link order, then +T over the 'to' ends */
        R* delta = w->node_tmp;
        memset(delta, 0, sizeof(R) * (size_t)N);
        for (int64_t l = 0; l < L; ++l) delta[t->lfrom[l]] = delta[t->lfrom[l]] + (-factor_T[l]);
        for (int64_t l = 0; l < L; ++l) delta[t->lto[l]] = delta[t->lto[l]] + factor_T[l];
        for (int64_t n = 0; n < N; ++n) V[n] = V[n] + delta[n];
        const R dts = vmax(dt, (R)1.0e-12);
        for (int64_t l = 0; l < L; ++l) Q[l] = factor_T[l] / dts;
        *transfer_abs = tabs;
    }
    double outfall_sum = 0.0;
    for (int64_t n = 0; n < N; ++n) V[n] = vmax(V[n], zero);
    for (int64_t n = 0; n < N; ++n) {
        const R head = node_head(t, V, n);
        const R dh = head - t->otw[n];
        R oq = t->ocoef[n] * t->oarea[n] * SQRT((R)(2.0 * GRAVITY) * FABS(dh)) * sgn(dh);
        oq = vmin(oq, vmax(t->ocap[n], zero));
        const R outfall = vmin(V[n], vmax(oq, zero) * dt);
        V[n] = vmax(V[n] - outfall, zero);
        outfall_sum += (double)outfall;
    }
    return outfall_sum;
}

/* v1_parity.py:371-404 */
static void return_node_overflow(const Topo* t, Work* w, R* h, R* V, double* over_out, double* unret_out) {
    const R zero = (R)0;
    const int64_t N = t->N, C = t->C, S = t->S;
    *over_out = 0.0;
    *unret_out = 0.0;
    if (N == 0) return;
    R* capacity = w->node_tmp2;
    R* overflow = w->head;
    for (int64_t n = 0; n < N; ++n) {
        capacity[n] = t->As[n] * vmax(t->dmax[n], zero);
        overflow[n] = vmax(V[n] - capacity[n], zero);
    }
    R* weight_sum = w->node_tmp;
    memset(weight_sum, 0, sizeof(R) * (size_t)N);
    for (int64_t k = 0; k < S; ++k) {
        const int valid = t->snode[k] >= 0 && t->scell[k] >= 0;
        const int32_t sn = valid ? t->snode[k] : 0;
        const R raw = valid ? vmax(t->sw[k], zero) : zero;
        weight_sum[sn] = weight_sum[sn] + raw;
    }
    R* returned_entry = w->entry_a;
    for (int64_t k = 0; k < S; ++k) {
        const int valid = t->snode[k] >= 0 && t->scell[k] >= 0;
        const int32_t sn = valid ? t->snode[k] : 0;
        const R raw = valid ? vmax(t->sw[k], zero) : zero;
        const R normalized = t->snorm ? raw : (valid ? raw / vmax(weight_sum[sn], (R)1.0e-12) : zero);
        returned_entry[k] = overflow[sn] * normalized;
    }
    memset(w->cell_tmp, 0, sizeof(R) * (size_t)C);
    for (int64_t k = 0; k < S; ++k) {
        const int valid = t->snode[k] >= 0 && t->scell[k] >= 0;
        const int32_t sc = valid ? t->scell[k] : 0;
        w->cell_tmp[sc] = w->cell_tmp[sc] + returned_entry[k] / t->area[sc];
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < C; ++c) h[c] = h[c] + w->cell_tmp[c];
    R* returned_node = w->node_tmp;      /* weight_sum no longer needed */
    memset(returned_node, 0, sizeof(R) * (size_t)N);
    uint8_t* explicit_ = w->link_pos;    /* sized max(L, N) */
    memset(explicit_, 0, (size_t)N);
    for (int64_t k = 0; k < S; ++k) {
        const int valid = t->snode[k] >= 0 && t->scell[k] >= 0;
        const int32_t sn = valid ? t->snode[k] : 0;
        returned_node[sn] = returned_node[sn] + returned_entry[k];
        if (valid) explicit_[sn] = 1;
    }
    memset(w->cell_tmp, 0, sizeof(R) * (size_t)C);
    double total = 0.0, unret = 0.0;
    for (int64_t n = 0; n < N; ++n) {
        const R rn = vmin(returned_node[n], overflow[n]);
        const R vn = V[n] - rn;
        const R residual = vmax(vn - capacity[n], zero);
        const int fb_valid = !explicit_[n] && t->node_cell[n] >= 0;
        const int32_t fcell = fb_valid ? t->node_cell[n] : 0;
        const R fallback = fb_valid ? residual : zero;
        w->cell_tmp[fcell] = w->cell_tmp[fcell] + fallback / t->area[fcell];
        V[n] = vn - fallback;
        total += (double)(rn + fallback);
        unret += (double)vmax(V[n] - capacity[n], zero);
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < C; ++c) h[c] = h[c] + w->cell_tmp[c];
    *over_out = total;
    *unret_out = unret;
}

/* v1_parity.py:415-427 */
static void recharge(const Topo* t, Work* w, R* h, R* storage, R dt, double* cap_out, double* loss_out) {
    const R zero = (R)0;
    *cap_out = 0.0;
    *loss_out = 0.0;
    if (t->Rr == 0) return;
    R* capture = w->entry_a;
    memset(w->cell_tmp, 0, sizeof(R) * (size_t)t->C);
    double cs = 0.0, ls = 0.0;
    for (int64_t r = 0; r < t->Rr; ++r) {
        const int32_t c = t->rcell[r];
        const R area = vmax(t->area[c], (R)1.0e-6);
        const R available = vmax(h[c] - t->rstop[r], zero) * area;
        capture[r] = vmin(vmin(available, t->rcc[r] * dt), vmax(t->rS[r] - storage[r], zero));
        w->cell_tmp[c] = w->cell_tmp[c] + capture[r];
        cs += (double)capture[r];
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) h[c] = vmax(h[c] - w->cell_tmp[c] / t->area[c], zero);
    for (int64_t r = 0; r < t->Rr; ++r) {
        const R st = storage[r] + capture[r];
        const R loss = vmin(st, t->rrc[r] * dt);
        storage[r] = st - loss;
        ls += (double)loss;
    }
    *cap_out = cs;
    *loss_out = ls;
}

/* Scratch sizes in elements of REAL (the caller allocates each buffer). */
void API(oracle_work_sizes)(const Topo* t, int64_t* out) {
    const int64_t nl = t->L > t->N ? t->L : t->N;
    const int64_t ns = t->S > t->Rr ? t->S : t->Rr;
    out[0] = t->E;            /* fl0..fr2 (6 of these), and bpart (doubles) */
    out[1] = t->C;            /* cell_tmp */
    out[2] = t->N;            /* node_tmp, node_tmp2, head */
    out[3] = t->I;            /* inlet_a, inlet_b */
    out[4] = t->L;            /* link_a, link_b */
    out[5] = ns;              /* entry_a */
    out[6] = nl;              /* link_pos (bytes) */
}

/* v1_parity.py:430-469: one step in the V1 process order.  State arrays are
 * updated in place; diag gets the 13 MassLedger terms of the step. */
int API(oracle_step)(const Topo* t, Work* w, R* h, R* hu, R* hv, R* wet_time, R* V, R* Q, R* storage,
                     double dt_in, double rain_rate, const R* ext_q, double max_speed, double capture_q,
                     double slot_wave_speed, double* diag) {
    const R dt = (R)dt_in;
    double s3[3];
    if (t->li) li_surface_step(t, w, h, hu, hv, dt, s3);
    else surface_flux_step(t, w, h, hu, hv, dt, s3);
    double rain, inf;
    sources_and_clamp(t, w, h, hu, hv, wet_time, dt, rain_rate, max_speed, &rain, &inf);
    const double ext = external_inflow(t, w, h, ext_q, dt);
    double cap = 0.0, sur = 0.0;
    if (t->I > 0) inlet_exchange(t, w, h, V, dt, capture_q, &cap, &sur);
    double tabs;
    const double outfall = pipe_links_and_outfalls(t, w, V, Q, dt, slot_wave_speed, &tabs);
    double over, unret;
    return_node_overflow(t, w, h, V, &over, &unret);
    double rcap, rloss;
    recharge(t, w, h, storage, dt, &rcap, &rloss);
    const R eps = (R)MIN_DEPTH, zero = (R)0;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < t->C; ++c) wet_time[c] = wet_time[c] + (h[c] > eps ? dt : zero);
    const double d[13] = {rain, inf, s3[0], s3[1], s3[2], ext, cap, sur, outfall, over, unret, rcap, rloss};
    memcpy(diag, d, sizeof(d));
    return 0;
}

int API(oracle_threads)(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
