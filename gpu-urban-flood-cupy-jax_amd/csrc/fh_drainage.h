// Drainage kernels: external inflow, inlet exchange, pipe links, node
// continuity (donor-limited transfers, outfalls, overflow) and the return /
// recharge pass, all as gathers over CSR tables built once on the host, so
// every sum has a fixed order and repeat runs are bit-identical.
//
// Semantics: hybrid_flood/jax_solver/v1_parity.py (relative to
// HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/):
//   _limit_by_group :259-263, _node_head :266-267, inlet_exchange :270-302,
//   _circular_area/_circular_radius :305-321, pipe_links_and_outfalls :324-368,
//   return_node_overflow :371-404, apply_external_inflow :407-412,
//   apply_recharge :415-427.  Each reference scatter_add becomes a gather whose
//   summation order equals the reference's sequential scatter order.
#pragma once
#include <vector>

#include "fh_device.h"

namespace fh {

template <typename real>
struct DrainDev {
    int N, I, Ci, L, S, Cr, R, Rcv, Cx;   // padded totals of the batch
    Seg seg_node, seg_link, seg_icell, seg_ret, seg_ext;
    // nodes
    const real *inv, *As, *dmax, *o_area, *o_coef, *o_tw, *o_cap;
    const int* nsc;
    real* V;
    real* Vp;               // node volume after the inlet exchange of this step
    real *sur_factor, *overflow, *fallback, *wsum;
    const int* explicit_;  // node has valid entry-table rows
    // node -> inlets, ascending inlet ("slots"): the inlet's cell and surcharge
    // limit, and per step the scaled surcharge and net exchange (capture - surcharge)
    const int *ni_ptr, *ni_cell;
    const real* ni_surq;
    real *ni_sur, *ni_net;
    const int *nl_ptr, *nl_ent;        // node -> incident links, ascending link, (l << 1) | is_to
    // node -> incident links, 'from' entries then 'to' entries (each ascending), with the far node
    const int *nl2_ptr, *nl2_ent, *nl2_oth;
    const int* ne_ptr;                 // node -> valid entries
    const real* ne_sw;                 //   their weights
    // inlet cells -> inlets (ascending): node, capture limit, node slot
    const int *ic_cells, *ic_ptr, *ic_node, *ic_slot;
    const real* ic_capq;
    // Row-aligned copies of each row's first table entries (0 where the row has
    // none), loaded together with the row itself rather than through its CSR
    // pointer, which saves one dependent memory round trip per kernel.
    const int* nf_cell;              // [N] first inlet of the node: cell, surcharge limit
    const real* nf_surq;
    const int *icf_node, *icf_slot;  // [inlet cells] first inlet: node, node slot, capture limit,
    const real *icf_capq, *icf_inv, *icf_As;   //   node invert and storage area
    const int *lk_a0, *lk_a1, *lk_b0, *lk_b1;  // [L] inlet-slot ranges of the two end nodes,
    const real *lk_inva, *lk_Asa, *lk_invb, *lk_Asb;   //   their inverts and storage areas
    const int* nle;                  // [KB][N] first KB entries of nl_ent
    const int *nl2e, *nl2o;          // [KB][N] first KB entries of nl2_ent / nl2_oth
    const real* nesw0;               // [N] first entry weight
    const int* rcf_node;             // [return cells] first entry: node, weight; first fallback node
    const real* rcf_sw;
    const int* rcf_fb;
    const real *ic_area, *rc_area;   // cell areas of the inlet cells and of the return cells, in list order
    // links
    const int *from, *to, *geom;
    const real *len, *man, *lcap, *linv, *D, *minor, *rw, *rh;
    real *Q, *req;
    // entries
    const int *snode;
    const real* sw;
    const int* snorm;        // [members] surcharge weights already normalized
    // return cells: entries (node, weight), fallback nodes, recharge rows
    const int *rc_cells, *rc_eptr, *rc_enode, *rc_fptr, *rc_fidx, *rc_rptr, *rc_ridx;
    const real* rc_esw;
    // recharge rows
    const real *rS, *rcc, *rrc, *rstop;
    real* rstore;
    // external
    const float* ext_q;      // per member [T_m][R_m], concatenated
    const int* rcv_qbase;    // [receivers] offset of the receiver's column in ext_q
    const int* rcv_stride;   // [receivers] R_m of the receiver's member
    const float* ext_q_override;
    int use_override;
    const int *xc_cells, *xc_ptr, *xc_idx;
};

template <typename real>
__device__ __forceinline__ real head_of(real inv, real As, real vol) {
    return inv + vmax(vol, real(0)) / vmax(As, real(1.0e-12));
}

template <typename real>
__device__ __forceinline__ real node_head(const DrainDev<real>& d, int n, real vol) {
    return head_of(d.inv[n], d.As[n], vol);
}

template <typename real>
__device__ __forceinline__ real orifice_q(const KP<real>& p, real dh) {
    const real sgn = dh > real(0) ? real(1) : (dh < real(0) ? real(-1) : real(0));
    return p.capture_q * sqrt(p.two_g * fabs(dh)) * sgn;
}

// Index of an entry loaded ahead of its row test: the entry itself when the
// row has it, else entry 0 (every table holds at least one element); values
// loaded for a missing entry are never used.
__device__ __forceinline__ int ahead(int k, int k1) { return k < k1 ? k : 0; }

// entries per row kept in the row-aligned copies of the node->link tables
constexpr int kRowEntries = 4;

// commit helper used by the last block of each ledger kernel
template <int NT>
__device__ __forceinline__ void commit_terms(Ctl* ctl, int counter, const double* part, int nparts,
                                             const int (&terms)[NT], double* lds) {
    double s[NT];
    sum_partials<NT>(part, nparts, s, lds);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&ctl->counters[counter], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            ctl->diag[terms[k]] = s[k];
            ctl->ledger[terms[k]] += s[k];
        }
    }
}

// v1_parity.py:407-412 -- h[recv] += max(q, 0) dt / A, q interpolated at t + dt/2.
template <typename real>
__global__ void __launch_bounds__(256) k_external(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
                                                 typename V4<real>::type* __restrict__ state,
                                                 const real* __restrict__ area, double* __restrict__ part) {
    const int b = blockIdx.x;
    const int mem = d.seg_ext.blk_member[b];
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    const int j = b * blockDim.x + threadIdx.x;
    const real dt = (real)ctl->dt_real;
    double added_sum = 0.0;
    if (j < d.seg_ext.mem_end[mem]) {
        const int c = d.xc_cells[j];
        const real A = area[c];
        const int lo = ctl->ext_lower, up = ctl->ext_upper;
        const real frac = (real)(float)ctl->ext_frac;
        real acc = real(0);
        for (int k = d.xc_ptr[j]; k < d.xc_ptr[j + 1]; ++k) {
            const int r = d.xc_idx[k];
            real q;
            if (d.use_override) {
                q = (real)d.ext_q_override[r];
            } else {
                const real ql = (real)d.ext_q[(size_t)d.rcv_qbase[r] + (size_t)lo * d.rcv_stride[r]];
                const real qu = (real)d.ext_q[(size_t)d.rcv_qbase[r] + (size_t)up * d.rcv_stride[r]];
                q = ql + frac * (qu - ql);
            }
            const real added = vmax(q, real(0)) * dt;
            acc = acc + added / A;
            added_sum += (double)added;
        }
        state[c].x = state[c].x + acc;
    }
    __shared__ double lds[16];
    double v[1] = {added_sum};
    block_sum<1>(v, lds);
    accumulate_slot<1>(part, b, v);
}

// The kernels below load what does not depend on the member's control block
// (the thread's own index rows) before testing whether the member's step is
// active: the grid covers exactly the padded index space, so those loads stay
// in bounds, and the gather chains start one memory round trip earlier.

// Inlet exchange, pass A (one thread per node): surcharge requests of the node's
// inlets and the node's donor factor  min(1, V / sum(requests)).
template <typename real>
__global__ void __launch_bounds__(256) k_inlet_node_factor(const Ctl* __restrict__ ctls, KP<real> p,
        DrainDev<real> d, const typename V4<real>::type* __restrict__ state) {
    typedef typename V4<real>::type V;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = d.seg_node.blk_member[blockIdx.x];
    const int k0 = d.ni_ptr[n], k1 = d.ni_ptr[n + 1];
    if (k0 == k1) return;
    const real vol = d.V[n];
    const real inv = d.inv[n], As = d.As[n];
    const real sq0 = d.nf_surq[n];                  // first inlet, loaded ahead
    const V s0 = state[d.nf_cell[n]];
    FH_ISSUE_LOADS_OF("s"(state));
    const Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    const real dt = (real)ctl->dt_real;
    const real nh = head_of(inv, As, vol);
    auto request = [&](const V& s, real sq) {
        const real q = orifice_q(p, (s.w + s.x) - nh);
        return vmin(vmax(-q, real(0)) * dt, vmax(sq, real(0)) * dt);
    };
    real total = real(0);
    {
        const real rs = request(s0, sq0);
        d.ni_sur[k0] = rs;
        total = total + rs;
    }
    for (int k = k0 + 1; k < k1; ++k) {
        const real rs = request(state[d.ni_cell[k]], d.ni_surq[k]);
        d.ni_sur[k] = rs;
        total = total + rs;
    }
    const real factor = vmin(real(1), vol / vmax(total, real(1.0e-12)));
    for (int k = k0; k < k1; ++k) d.ni_sur[k] = d.ni_sur[k] * factor;
}

// Inlet exchange, pass B (one thread per inlet cell): capture requests, the
// cell's donor factor min(1, h A / sum), captured volumes, and the depth update
// h = max(h + sum((sur - cap) / A), 0).  Each inlet's net exchange
// (capture - surcharge) goes to its node slot for the node volume update.
template <typename real>
__global__ void __launch_bounds__(256) k_inlet_cell(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
        typename V4<real>::type* __restrict__ state, const real* __restrict__ area, double* __restrict__ part) {
    typedef typename V4<real>::type V;
    const int b = blockIdx.x;
    const int j = b * blockDim.x + threadIdx.x;
    const int mem = d.seg_icell.blk_member[b];
    const int c = d.ic_cells[j];
    const int k0 = d.ic_ptr[j], k1 = d.ic_ptr[j + 1];
    V s = state[c];
    const real A = d.ic_area[j];
    // first inlet of the cell, loaded ahead (rows without one read node 0 / slot 0, never used)
    const int n0 = d.icf_node[j], slot0 = d.icf_slot[j];
    const real capq0 = d.icf_capq[j], inv0 = d.icf_inv[j], As0 = d.icf_As[j];
    const real V0 = d.V[n0], sur0 = d.ni_sur[slot0];
    SlotAcc<2> acc2;
    acc2.prefetch(part, b);
    FH_ISSUE_LOADS_OF("s"(state), "s"(area), "s"(part));
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    double cap_sum = 0.0, sur_sum = 0.0;
    if (j < d.seg_icell.mem_end[mem]) {
        const real dt = (real)ctl->dt_real;
        const real sh = s.w + s.x;
        real total = real(0);
        real rc0 = real(0);
        for (int k = k0; k < k1; ++k) {
            real nh;
            real cq;
            if (k == k0) {
                nh = head_of(inv0, As0, V0);
                cq = capq0;
            } else {
                const int n = d.ic_node[k];
                nh = head_of(d.inv[n], d.As[n], d.V[n]);
                cq = d.ic_capq[k];
            }
            const real q = orifice_q(p, sh - nh);
            const real rc = vmin(vmax(q, real(0)) * dt, vmax(cq, real(0)) * dt);
            if (k == k0) rc0 = rc;
            else d.ni_net[d.ic_slot[k]] = rc;          // request, scaled below
            total = total + rc;
        }
        const real factor = vmin(real(1), (s.x * A) / vmax(total, real(1.0e-12)));
        real acc = real(0);
        for (int k = k0; k < k1; ++k) {
            const int slot = k == k0 ? slot0 : d.ic_slot[k];
            const real cp = (k == k0 ? rc0 : d.ni_net[slot]) * factor;
            const real sr = k == k0 ? sur0 : d.ni_sur[slot];
            d.ni_net[slot] = cp - sr;
            acc = acc + (sr - cp) / A;
            cap_sum += (double)cp;
            sur_sum += (double)sr;
        }
        s.x = vmax(s.x + acc, real(0));
        state[c].x = s.x;
    }
    __shared__ double lds[32];
    double v[2] = {cap_sum, sur_sum};
    block_sum<2>(v, lds);
    acc2.commit(part, b, v);
}

// Node volume after the inlet exchange (v1_parity.py:292-298):
// max(V + sum over the node's inlets of (capture - surcharge), 0), for every node.
// net0 is ni_net[k0], loaded ahead by the caller (unused when the node has no inlet).
template <typename real>
__device__ __forceinline__ real post_inlet_volume(const DrainDev<real>& d, int k0, int k1, real vol, real net0) {
    real acc = real(0);
    for (int k = k0; k < k1; ++k) acc = acc + (k == k0 ? net0 : d.ni_net[k]);
    return vmax(vol + acc, real(0));
}

template <typename real>
__device__ __forceinline__ void circular_theta(real depth, real diameter, real& radius, real& theta) {
    radius = real(0.5) * diameter;
    const real od = vmin(vmax(depth, real(0)), diameter);
    real x = (radius - od) / radius;
    x = vmin(vmax(x, real(-1)), real(1));
    theta = real(2) * acos(x);
}

// Static data of one link, loaded up front.
template <typename real>
struct LinkData {
    real linv, D, len, man, q0, minor, lcap, rw, rh;
    int geom;
};

template <typename real>
__device__ __forceinline__ LinkData<real> load_link(const DrainDev<real>& d, int l) {
    LinkData<real> k;
    k.linv = d.linv[l]; k.D = d.D[l]; k.len = d.len[l]; k.man = d.man[l]; k.q0 = d.Q[l];
    k.minor = d.minor[l]; k.lcap = d.lcap[l]; k.rw = d.rw[l]; k.rh = d.rh[l]; k.geom = d.geom[l];
    return k;
}

// Pipe link momentum (v1_parity.py:324-347): semi-implicit Saint-Venant update
// with Manning friction and minor losses, clipped to capacity; returns the
// signed requested transfer Q dt.  ha / hb are the end-node heads.
template <typename real>
__device__ __forceinline__ real link_request(const KP<real>& p, const LinkData<real>& k, real ha, real hb,
                                             real dt) {
    const real zero = real(0), g = p.g;
    const real depth = vmax(real(0.5) * (ha + hb) - k.linv, zero);
    const real Draw = k.D;
    const real Dm = vmax(Draw, real(1.0e-12));
    const real slot_width = g * (real(0.25 * M_PI) * (Dm * Dm)) / p.slot_coef;
    real radius_c, theta;
    circular_theta(depth, Dm, radius_c, theta);
    const real open_c = real(0.5) * (radius_c * radius_c) * (theta - sin(theta));
    const real full = real(0.25 * M_PI) * (Dm * Dm);
    const real circ_area = depth >= Dm ? full : open_c;
    real flow_area = circ_area + slot_width * vmax(depth - Draw, zero);
    const real perim = depth >= Dm ? real(M_PI) * Dm : radius_c * theta;
    real radius = circ_area / vmax(perim, real(1.0e-12));
    if (k.geom == 2) {
        const real od = vmin(vmax(depth, zero), vmax(k.rh, real(1.0e-6)));
        const real oa = vmax(k.rw, real(1.0e-6)) * od;
        flow_area = oa;
        radius = oa / vmax(k.rw + real(2) * od, real(1.0e-12));
    }
    const real length = vmax(k.len, real(1.0e-6));
    const real area = vmax(flow_area, real(1.0e-12));
    radius = vmax(radius, real(1.0e-12));
    const real slope = (ha - hb) / length;
    const real nm = vmax(k.man, zero);
    const real absq = fabs(k.q0);
    const real friction = g * (nm * nm) * absq / (pow(radius, p.pow43) * area);
    const real minor = vmax(k.minor, zero) * absq / (real(2) * length * area);
    real cand = (k.q0 + g * area * dt * slope) / (real(1) + dt * (friction + minor));
    const real cap = vmax(k.lcap, zero);
    cand = vmin(vmax(cand, -cap), cap);
    return cand * dt;
}

// Pipe links (one thread per link): the requested transfer Q dt from the
// post-inlet volumes of the link's two end nodes (v1_parity.py:324-347).
template <typename real>
__global__ void __launch_bounds__(256) k_link(const Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = d.seg_link.blk_member[blockIdx.x];
    const int a = d.from[l], b = d.to[l];
    const LinkData<real> lk = load_link(d, l);
    const int a0 = d.lk_a0[l], a1 = d.lk_a1[l], b0 = d.lk_b0[l], b1 = d.lk_b1[l];
    const real inva = d.lk_inva[l], Asa = d.lk_Asa[l], invb = d.lk_invb[l], Asb = d.lk_Asb[l];
    const real Va = d.V[a], Vb = d.V[b];
    const real neta = d.ni_net[ahead(a0, a1)], netb = d.ni_net[ahead(b0, b1)];
    FH_ISSUE_LOADS();
    const Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    if (l >= d.seg_link.mem_end[mem]) return;
    const real dt = (real)ctl->dt_real;
    const real va = post_inlet_volume(d, a0, a1, Va, neta);
    const real vb = a == b ? va : post_inlet_volume(d, b0, b1, Vb, netb);
    d.req[l] = link_request(p, lk, head_of(inva, Asa, va), head_of(invb, Asb, vb), dt);
}

// Donor factors (one thread per node): the node's post-inlet volume and
// min(1, V / sum of |requested| over the links it donates to, in ascending link
// order) -- v1_parity.py:348-352.
template <typename real>
__global__ void __launch_bounds__(256) k_node_factor(const Ctl* __restrict__ ctls, DrainDev<real> d) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = d.seg_node.blk_member[blockIdx.x];
    const int i0 = d.ni_ptr[n], i1 = d.ni_ptr[n + 1];
    const int k0 = d.nl_ptr[n], k1 = d.nl_ptr[n + 1];
    const real vol = d.V[n];
    // the first inlet slot and the first KB link entries, loaded ahead
    constexpr int KB = kRowEntries;
    const int N = d.N;
    const real net0 = d.ni_net[ahead(i0, i1)];
    int en[KB];
    real r[KB];
#pragma unroll
    for (int u = 0; u < KB; ++u) en[u] = d.nle[u * N + n];
#pragma unroll
    for (int u = 0; u < KB; ++u) r[u] = d.req[en[u] >> 1];
    FH_ISSUE_LOADS();
    if (!ctls[mem].active) return;
    if (n >= d.seg_node.mem_end[mem]) return;
    const real vn = post_inlet_volume(d, i0, i1, vol, net0);
    d.Vp[n] = vn;
    real total = real(0);
    for (int kb = k0; kb < k1; kb += KB) {
        if (kb != k0) {
#pragma unroll
            for (int u = 0; u < KB; ++u) en[u] = kb + u < k1 ? d.nl_ent[kb + u] : 0;
#pragma unroll
            for (int u = 0; u < KB; ++u) r[u] = kb + u < k1 ? d.req[en[u] >> 1] : real(0);
        }
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            if (kb + u >= k1) break;
            const bool is_to = en[u] & 1;
            const bool positive = r[u] >= real(0);
            if (positive != is_to) total = total + fabs(r[u]);   // donor = from if positive else to
        }
    }
    d.sur_factor[n] = vmin(real(1), vn / vmax(total, real(1.0e-12)));
}

// Node continuity (one thread per node): donor-limited link transfers
// (reference scatter order: all 'from' links, then all 'to' links), outfall,
// overflow above storage capacity (weighted entry table, then the node surface
// cell fallback); writes the link flow for links this node is 'from' of.
template <typename real>
__global__ void __launch_bounds__(256) k_node_update(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
                                                    double* __restrict__ part) {
    const int b = blockIdx.x;
    const int n = b * blockDim.x + threadIdx.x;
    const int mem = d.seg_node.blk_member[b];
    const real zero = real(0);
    const int k0 = d.nl2_ptr[n], k1 = d.nl2_ptr[n + 1];
    const int e0 = d.ne_ptr[n], e1 = d.ne_ptr[n + 1];
    const real vp = d.Vp[n], f_self = d.sur_factor[n];
    const real o_tw = d.o_tw[n], o_coef = d.o_coef[n], o_area = d.o_area[n], o_cap = d.o_cap[n];
    const real inv = d.inv[n], As = d.As[n], dmax = d.dmax[n];
    const int expl = d.explicit_[n], nsc = d.nsc[n];
    // the first KB link entries and the first entry weight, loaded ahead
    constexpr int KB = kRowEntries;
    const int N = d.N;
    int en[KB], ot[KB];
    real r[KB], fo[KB];
#pragma unroll
    for (int u = 0; u < KB; ++u) {
        en[u] = d.nl2e[u * N + n];
        ot[u] = d.nl2o[u * N + n];
    }
#pragma unroll
    for (int u = 0; u < KB; ++u) {
        r[u] = d.req[en[u] >> 1];
        fo[u] = d.sur_factor[ot[u]];
    }
    const real sw0 = d.nesw0[n];
    SlotAcc<3> acc3;
    acc3.prefetch(part, b);
    FH_ISSUE_LOADS_OF("s"(part));
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    double out_sum = 0.0, total_sum = 0.0, unret_sum = 0.0;
    if (n < d.seg_node.mem_end[mem]) {
        const real dt = (real)ctl->dt_real;
        real vol = vp;
        {
            real delta = zero;
            for (int kb = k0; kb < k1; kb += KB) {
                if (kb != k0) {
#pragma unroll
                    for (int u = 0; u < KB; ++u) en[u] = kb + u < k1 ? d.nl2_ent[kb + u] : 0;
#pragma unroll
                    for (int u = 0; u < KB; ++u) {
                        r[u] = kb + u < k1 ? d.req[en[u] >> 1] : zero;
                        fo[u] = kb + u < k1 ? d.sur_factor[d.nl2_oth[kb + u]] : zero;
                    }
                }
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    if (kb + u >= k1) break;
                    const int is_to = en[u] & 1;
                    const bool positive = r[u] >= zero;
                    // donor = from if positive else to; this node is 'to' when is_to
                    const real f = (positive != (is_to != 0)) ? f_self : fo[u];
                    const real T = (positive ? real(1) : real(-1)) * fabs(r[u]) * f;
                    if (!is_to) {
                        delta = delta + (-T);
                        d.Q[en[u] >> 1] = T / vmax(dt, real(1.0e-12));
                    } else {
                        delta = delta + T;
                    }
                }
            }
            vol = vmax(vol + delta, zero);
        }
        // outfall (v1_parity.py:361-367)
        const real hd = head_of(inv, As, vol) - o_tw;
        const real sgn = hd > zero ? real(1) : (hd < zero ? real(-1) : zero);
        real oq = o_coef * o_area * sqrt(p.two_g * fabs(hd)) * sgn;
        oq = vmin(oq, vmax(o_cap, zero));
        const real outf = vmin(vol, vmax(oq, zero) * dt);
        vol = vmax(vol - outf, zero);
        // overflow (v1_parity.py:371-404)
        const real capacity = As * vmax(dmax, zero);
        const real over = vmax(vol - capacity, zero);
        real returned = zero;
        real wsum = zero;
        if (d.S > 0) {
            for (int k = e0; k < e1; ++k) wsum = wsum + vmax(k == e0 ? sw0 : d.ne_sw[k], zero);
            for (int k = e0; k < e1; ++k) {
                const real raw = vmax(k == e0 ? sw0 : d.ne_sw[k], zero);
                const real norm = d.snorm[mem] ? raw : raw / vmax(wsum, real(1.0e-12));
                returned = returned + over * norm;
            }
            returned = vmin(returned, over);
        }
        vol = vol - returned;
        const real residual = vmax(vol - capacity, zero);
        const real fb = (!expl && nsc >= 0) ? residual : zero;
        vol = vol - fb;
        d.V[n] = vol;
        d.overflow[n] = over;
        d.wsum[n] = wsum;
        d.fallback[n] = fb;
        out_sum = (double)outf;
        total_sum = (double)(returned + fb);
        unret_sum = (double)vmax(vol - capacity, zero);
    }
    __shared__ double lds[48];
    double v[3] = {out_sum, total_sum, unret_sum};
    block_sum<3>(v, lds);
    acc3.commit(part, b, v);
}

// Return pass (one thread per receiving cell): weighted entry returns, node
// fallback returns, then recharge capture/loss (v1_parity.py:415-427).
template <typename real>
__global__ void __launch_bounds__(256) k_return(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
        typename V4<real>::type* __restrict__ state, const real* __restrict__ area, double* __restrict__ part) {
    const int b = blockIdx.x;
    const int j = b * blockDim.x + threadIdx.x;
    const int mem = d.seg_ret.blk_member[b];
    const real zero = real(0);
    const int c = d.rc_cells[j];
    const int ke0 = d.rc_eptr[j], ke1 = d.rc_eptr[j + 1];
    const int kf0 = d.rc_fptr[j], kf1 = d.rc_fptr[j + 1];
    const real A = d.rc_area[j];
    const real h0 = state[c].x;
    // first entry and first fallback node, loaded ahead
    const int ne0 = d.rcf_node[j];
    const real esw0 = d.rcf_sw[j];
    const real over0 = d.overflow[ne0], wsum0 = d.wsum[ne0];
    const real fb0 = d.fallback[d.rcf_fb[j]];
    SlotAcc<2> acc2;
    acc2.prefetch(part, b);
    FH_ISSUE_LOADS_OF("s"(state), "s"(area), "s"(part));
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    double cap_sum = 0.0, loss_sum = 0.0;
    if (j < d.seg_ret.mem_end[mem]) {
        const real dt = (real)ctl->dt_real;
        real h = h0;
        if (d.S > 0) {
            real acc = zero;
            for (int k = ke0; k < ke1; ++k) {
                real raw, ov, ws;
                if (k == ke0) {
                    raw = vmax(esw0, zero); ov = over0; ws = wsum0;
                } else {
                    const int n = d.rc_enode[k];
                    raw = vmax(d.rc_esw[k], zero); ov = d.overflow[n]; ws = d.wsum[n];
                }
                const real norm = d.snorm[mem] ? raw : raw / vmax(ws, real(1.0e-12));
                acc = acc + (ov * norm) / A;
            }
            h = h + acc;
        }
        {
            real acc = zero;
            for (int k = kf0; k < kf1; ++k) acc = acc + (k == kf0 ? fb0 : d.fallback[d.rc_fidx[k]]) / A;
            h = h + acc;
        }
        if (d.R > 0 && d.rc_rptr[j + 1] > d.rc_rptr[j]) {
            const real Ar = vmax(A, real(1.0e-6));
            real removed = zero;
            for (int k = d.rc_rptr[j]; k < d.rc_rptr[j + 1]; ++k) {
                const int r = d.rc_ridx[k];
                const real avail = vmax(h - d.rstop[r], zero) * Ar;
                const real s0 = d.rstore[r];
                const real capt = vmin(vmin(avail, d.rcc[r] * dt), vmax(d.rS[r] - s0, zero));
                removed = removed + capt;
                const real st = s0 + capt;
                const real loss = vmin(st, d.rrc[r] * dt);
                d.rstore[r] = st - loss;
                cap_sum += (double)capt;
                loss_sum += (double)loss;
            }
            h = vmax(h - removed / A, zero);
        }
        state[c].x = h;
    }
    __shared__ double lds[32];
    double v[2] = {cap_sum, loss_sum};
    block_sum<2>(v, lds);
    acc2.commit(part, b, v);
}

}  // namespace fh
