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
    // Return rows k_return handles (rx_rows, blocks seg_rx): those with entry
    // returns, recharge, or several fallback nodes.  Every other row has exactly
    // one fallback node and nothing else; that node's thread in k_node_update
    // applies it: nr_cell [N] (-1: none) and nr_area [N] its cell and area.
    Seg seg_rx;
    const int* rx_rows;
    const int* nr_cell;
    const real* nr_area;
    // nodes
    const real *inv, *As, *dmax;
    // [N] bit 0: the overflow left after the entry table falls back to the
    // node's surface cell (no valid entry rows, surface cell >= 0); bits 1..:
    // outfall slot + 1, or 0 where the outfall term is exactly zero (coefficient
    // x area == 0 with finite tailwater and capacity)
    const int* nflag;
    const typename V4<real>::type* outfall;   // [slots] {tailwater, coefficient, area, capacity}
    real* V;
    real* Vp;               // node volume after the inlet exchange of this step
    real *sur_factor, *overflow, *fallback, *wsum;
    int store_returns;      // k_return runs: k_node_update stores overflow, weight sum, fallback
    // node -> inlets, ascending inlet ("slots"): the inlet's cell and surcharge
    // limit, and per step the net exchange (capture - surcharge)
    const int *ni_ptr, *ni_cell;
    const real* ni_surq;
    real* ni_net;
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
    // per inlet entry (ic_* order): its cell, its node's invert and storage area,
    // and the inlets of its node ("siblings", in slot order): first slot, count,
    // and the first KB siblings' cells and surcharge limits ([KB][entries]; the
    // entry's own cell and 0 past the count)
    const int *ie_cell, *ie_row, *ie_s0, *ie_ns, *ie_sc;
    const real *ie_inv, *ie_As, *ie_sq;
    real* ic_h;                      // [inlet cells] depth after the inlet exchange of this step
    const int* nle;                  // [KB][N] first KB entries of nl_ent
    // Link work items of each block of kNodeBlock nodes (k_link_node): the links
    // whose 'from' node is in the block, then the links entering it from other
    // blocks, in the block's entry order; lb_ptr [node blocks + 1] delimits them.
    // Per item: link, end nodes and their inlet-slot ranges.  nl_src gives each
    // nl_ent entry the block-local item holding its link's request, nls [KB][N]
    // the same for the first KB entries of each node.
    const int *lb_ptr, *it_l, *it_a, *it_b, *it_a0, *it_a1, *it_b0, *it_b1;
    const int *nl_src, *nls;
    // static link data, two vectors per link: {invert, diameter, length, Manning n},
    // {minor loss, capacity, rectangle width, rectangle height}
    const typename V4<real>::type *lk_a, *lk_b;
    const int *nl2e, *nl2o;          // [KB][N] first KB entries of nl2_ent / nl2_oth
    const real* nesw0;               // [N] first entry weight
    const int* rcf_node;             // [return cells] first entry: node, weight; first fallback node
    const real* rcf_sw;
    const int* rcf_fb;
    const real *ic_area, *rc_area;   // cell areas of the inlet cells and of the return cells, in list order
    // links: geometry code (2: rectangle), flow, this step's requested transfer
    const int* geom;
    real *Q, *req;
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
    // dry-weather stepping (triangular_backend.py:833-952): no surface exchange, the
    // outfall tailwater optionally overridden per node, overflow above capacity
    // reported and (dry_cap) capped instead of returned
    int dry, dry_cap;
    const real* dry_tw;
};

template <typename real>
__device__ __forceinline__ real head_of(real inv, real As, real vol) {
    return inv + Divisor<real>(vmax(As, real(1.0e-12))).div(vmax(vol, real(0)));
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
// inlet cells per block of k_inlet (the inlet-cell layout's tile) and its threads.
// Small blocks spread the inlet exchange over every CU (the production network's
// 61,317 inlets fill about a thousand blocks): with 64 rows the step ran 0.5 % faster
// than with 128 and 1.2 % faster than with 256, the state bit for bit the same.
constexpr int kInletRows = 64;
constexpr int kInletThreads = 128;
// nodes per block of the node kernels (the node layout's tile), and threads
// per block of k_link_node (one link item each).  With 128 nodes the step ran
// 0.5 % faster than with 256 (the state bit for bit the same); with 64 it ran
// 1.2 % slower, as more links cross blocks and become items of both.
constexpr int kNodeBlock = 128;
constexpr int kLinkThreads = 256;

// v1_parity.py:407-412 -- h[recv] += max(q, 0) dt / A, q interpolated at t + dt/2.
template <typename real>
__global__ void __launch_bounds__(256) k_external(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
                                                 typename V4<real>::type* __restrict__ state,
                                                 const real* __restrict__ area, double* __restrict__ part) {
    const int b = xcd_block(blockIdx.x, gridDim.x, p.remap);
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
            acc = acc + Divisor<real>(A).div(added);
            added_sum += (double)added;
        }
        state[c].x = state[c].x + acc;
    }
    __shared__ double lds[16];
    double v[1] = {added_sum};
    block_sum<1>(v, lds);
    accumulate_slot<1>(part, b, v);
}

// Every drainage kernel takes its logical block through xcd_block (the XCD-aware
// order of the surface kernels): nodes, links, inlet and return cells are stored
// in cell order, so a contiguous range of blocks on one XCD keeps the gathers of
// neighbouring rows in that XCD's L2 (inlet 7.3 -> 6.6 us, node update 7.4 ->
// 6.8 us on the production workload).
//
// The kernels below load what does not depend on the member's control block
// (the thread's own index rows) before testing whether the member's step is
// active: the grid covers exactly the padded index space, so those loads stay
// in bounds, and the gather chains start one memory round trip earlier.

// Surcharge request of one inlet (v1_parity.py:275-283): min(max(-q, 0) dt,
// max(limit, 0) dt), q the orifice flow from the cell's water surface to the
// node head.
template <typename real>
__device__ __forceinline__ real surcharge_request(const KP<real>& p, const typename V4<real>::type& s, real sq,
                                                  real nh, real dt) {
    const real q = orifice_q(p, (s.w + s.x) - nh);
    return vmin(vmax(-q, real(0)) * dt, vmax(sq, real(0)) * dt);
}

// Inlet exchange, v1_parity.py:270-302, one block per kInletRows inlet cells
// with kInletThreads threads.  Each thread takes one inlet of the block's cells
// (the ic_* entries of those rows, contiguous): the inlet's surcharge request
// scaled by its node's donor factor min(1, V / sum of the requests of all the
// node's inlets, in slot order -- each recomputed from its cell), and its
// capture request, into LDS.  After the barrier one thread per cell forms the
// cell's donor factor min(1, h A / sum of its capture requests), the depth
// h = max(h + sum((sur - cap) / A), 0) (to ic_h: the state keeps the depth the
// sibling requests read until k_node_update applies it) and the ledger sums;
// after a second barrier each inlet stores its net exchange (capture -
// surcharge) in its node slot.  Dynamic LDS: two reals per inlet of the block
// and one per cell.
template <typename real, bool GATE>
__global__ void __launch_bounds__(kInletThreads) k_inlet(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
        const typename V4<real>::type* __restrict__ state, double* __restrict__ part) {
    typedef typename V4<real>::type V;
    constexpr int KB = kRowEntries;
    extern __shared__ __align__(16) unsigned char fh_lds_dyn[];
    const int b = xcd_block(blockIdx.x, gridDim.x, p.remap);
    const int t = threadIdx.x;
    const int mem = d.seg_icell.blk_member[b];
    if (GATE && !ctls[mem].active) return;      // the whole block: one member
    const int mend = d.seg_icell.mem_end[mem];
    const int nent = d.ic_ptr[d.Ci];              // inlet entries of the batch
    const int e0 = d.ic_ptr[b * kInletRows], e1 = d.ic_ptr[b * kInletRows + kInletRows];
    real* lsur = reinterpret_cast<real*>(fh_lds_dyn);
    real* lrc = lsur + (e1 - e0);
    real* lfac = lrc + (e1 - e0);
    // this thread's inlet, loaded ahead (threads past the block's inlets read
    // entry 0, never used)
    const int e = e0 + t < e1 ? e0 + t : 0;
    const int n = d.ic_node[e], slot = d.ic_slot[e];
    const real capq = d.ic_capq[e], inv = d.ie_inv[e], As = d.ie_As[e];
    const int sib0 = d.ie_s0[e], nsib = d.ie_ns[e];
    const V sc = state[d.ie_cell[e]];
    const real Vn = d.V[n];
    V ss[KB];
    real sq[KB];
#pragma unroll
    for (int u = 0; u < KB; ++u) {
        ss[u] = state[d.ie_sc[u * nent + e]];
        sq[u] = d.ie_sq[u * nent + e];
    }
    // this thread's cell (threads < kInletRows)
    const bool row_thread = t < kInletRows;
    const int j = b * kInletRows + (row_thread ? t : 0);
    const int k0 = d.ic_ptr[j], k1 = d.ic_ptr[j + 1];
    const real h0 = state[d.ic_cells[j]].x;
    const real A = d.ic_area[j];
    SlotAcc<2> acc2;
    acc2.prefetch(part, b);
    FH_ISSUE_LOADS_OF("s"(state), "s"(part));
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;                 // the whole block: one member
    const real dt = (real)ctl->dt_real;
    const bool live_e = e0 + t < e1;
    if (live_e) {
        const real nh = head_of(inv, As, Vn);
        real total = real(0), own = real(0);
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            if (u >= nsib) break;
            const real rs = surcharge_request(p, ss[u], sq[u], nh, dt);
            total = total + rs;
            if (sib0 + u == slot) own = rs;
        }
        for (int u = KB; u < nsib; ++u) {
            const real rs = surcharge_request(p, state[d.ni_cell[sib0 + u]], d.ni_surq[sib0 + u], nh, dt);
            total = total + rs;
            if (sib0 + u == slot) own = rs;
        }
        lsur[t] = own * vmin(real(1), Divisor<real>(vmax(total, real(1.0e-12))).div(Vn));
        const real q = orifice_q(p, (sc.w + sc.x) - nh);
        lrc[t] = vmin(vmax(q, real(0)) * dt, vmax(capq, real(0)) * dt);
    }
    for (int i = e0 + t + kInletThreads; i < e1; i += kInletThreads) {   // blocks with more inlets than threads
        const int ni = d.ic_node[i], si = d.ic_slot[i];
        const real Vi = d.V[ni];
        const real nh = head_of(d.ie_inv[i], d.ie_As[i], Vi);
        real total = real(0), own = real(0);
        for (int u = 0; u < d.ie_ns[i]; ++u) {
            const int k = d.ie_s0[i] + u;
            const real rs = surcharge_request(p, state[d.ni_cell[k]], d.ni_surq[k], nh, dt);
            total = total + rs;
            if (k == si) own = rs;
        }
        lsur[i - e0] = own * vmin(real(1), Divisor<real>(vmax(total, real(1.0e-12))).div(Vi));
        const V si_s = state[d.ie_cell[i]];
        const real q = orifice_q(p, (si_s.w + si_s.x) - nh);
        lrc[i - e0] = vmin(vmax(q, real(0)) * dt, vmax(d.ic_capq[i], real(0)) * dt);
    }
    __syncthreads();
    double cap_sum = 0.0, sur_sum = 0.0;
    if (row_thread && j < mend) {
        real total = real(0);
        for (int k = k0; k < k1; ++k) total = total + lrc[k - e0];
        const real factor = vmin(real(1), Divisor<real>(vmax(total, real(1.0e-12))).div(h0 * A));
        real acc = real(0);
        for (int k = k0; k < k1; ++k) {
            const real cp = lrc[k - e0] * factor;
            const real sr = lsur[k - e0];
            acc = acc + Divisor<real>(A).div(sr - cp);
            cap_sum += (double)cp;
            sur_sum += (double)sr;
        }
        lfac[t] = factor;
        d.ic_h[j] = vmax(h0 + acc, real(0));
    }
    __syncthreads();
    // each inlet's net exchange, with its cell's factor (the entries of row r
    // are ic_ptr[r] .. ic_ptr[r + 1], so the row is found from the entry)
    if (live_e) {
        const int r = d.ie_row[e] - b * kInletRows;
        d.ni_net[slot] = lrc[t] * lfac[r] - lsur[t];
    }
    for (int i = e0 + t + kInletThreads; i < e1; i += kInletThreads) {
        const int r = d.ie_row[i] - b * kInletRows;
        d.ni_net[d.ic_slot[i]] = lrc[i - e0] * lfac[r] - lsur[i - e0];
    }
    __shared__ double lds[32];
    double v[2] = {cap_sum, sur_sum};
    block_sum<2>(v, lds);
    acc2.commit(part, b, v);
}

// Node volume after the inlet exchange (v1_parity.py:292-298):
// max(V + sum over the node's inlets of (capture - surcharge), 0), for every node.
// net0, net1 are ni_net[k0], ni_net[k0 + 1], loaded ahead by the caller (unused
// where the node has fewer inlets); later slots are read here.
template <typename real>
__device__ __forceinline__ real post_inlet_volume(const DrainDev<real>& d, int k0, int k1, real vol, real net0,
                                                  real net1) {
    real acc = real(0);
    if (k0 < k1) acc = acc + net0;
    if (k0 + 1 < k1) acc = acc + net1;
    for (int k = k0 + 2; k < k1; ++k) acc = acc + d.ni_net[k];
    return vmax(vol + acc, real(0));
}

// The first two inlet-slot nets of a node with inlet slots [k0, k1), for
// post_inlet_volume (slot 0 where the node has fewer: never used).
template <typename real>
struct Nets2 {
    real n0, n1;
};
template <typename real>
__device__ __forceinline__ Nets2<real> first_nets(const DrainDev<real>& d, int k0, int k1) {
    return Nets2<real>{d.ni_net[ahead(k0, k1)], d.ni_net[ahead(k0 + 1, k1)]};
}

template <typename real>
__device__ __forceinline__ void circular_theta(real depth, real diameter, real& radius, real& theta) {
    radius = real(0.5) * diameter;
    const real od = vmin(vmax(depth, real(0)), diameter);
    real x = Divisor<real>(radius).div(radius - od);
    x = vmin(vmax(x, real(-1)), real(1));
    theta = real(2) * acos(x);
}

// Data of one link for its momentum update.
template <typename real>
struct LinkData {
    real linv, D, len, man, q0, minor, lcap, rw, rh;
    int geom;
};

template <typename real>
__device__ __forceinline__ LinkData<real> load_link(const DrainDev<real>& d, int l) {
    LinkData<real> k;
    const typename V4<real>::type a = d.lk_a[l], b = d.lk_b[l];
    k.linv = a.x; k.D = a.y; k.len = a.z; k.man = a.w;
    k.minor = b.x; k.lcap = b.y; k.rw = b.z; k.rh = b.w;
    k.q0 = d.Q[l]; k.geom = d.geom[l];
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
    const real slot_width = Divisor<real>(p.slot_coef).div(g * (real(0.25 * M_PI) * (Dm * Dm)));
    real radius_c, theta;
    circular_theta(depth, Dm, radius_c, theta);
    const real open_c = real(0.5) * (radius_c * radius_c) * (theta - sin(theta));
    const real full = real(0.25 * M_PI) * (Dm * Dm);
    const real circ_area = depth >= Dm ? full : open_c;
    real flow_area = circ_area + slot_width * vmax(depth - Draw, zero);
    const real perim = depth >= Dm ? real(M_PI) * Dm : radius_c * theta;
    real radius = Divisor<real>(vmax(perim, real(1.0e-12))).div(circ_area);
    if (k.geom == 2) {
        const real od = vmin(vmax(depth, zero), vmax(k.rh, real(1.0e-6)));
        const real oa = vmax(k.rw, real(1.0e-6)) * od;
        flow_area = oa;
        radius = Divisor<real>(vmax(k.rw + real(2) * od, real(1.0e-12))).div(oa);
    }
    const real length = vmax(k.len, real(1.0e-6));
    const real area = vmax(flow_area, real(1.0e-12));
    radius = vmax(radius, real(1.0e-12));
    const real slope = Divisor<real>(length).div(ha - hb);
    const real nm = vmax(k.man, zero);
    const real absq = fabs(k.q0);
    const real friction = Divisor<real>(fric_pow(radius, p.pow43_dln2) * area).div(g * (nm * nm) * absq);
    const real minor = Divisor<real>(real(2) * length * area).div(vmax(k.minor, zero) * absq);
    real cand = Divisor<real>(real(1) + dt * (friction + minor)).div(k.q0 + g * area * dt * slope);
    const real cap = vmax(k.lcap, zero);
    cand = vmin(vmax(cand, -cap), cap);
    return cand * dt;
}

// Pipe links and donor factors, v1_parity.py:324-352, one block per
// kNodeBlock nodes.  Every thread computes the requested transfer Q dt of one
// of the block's link items (from the post-inlet volumes of both end nodes)
// into LDS; after the block barrier the first kNodeBlock threads, one per node,
// compute the node's post-inlet volume and donor factor min(1, V / sum of
// |requested| over the links it donates to, in ascending link order), reading
// each request from its item.  A link inside the block is one item; a link
// between two blocks is an item of both, which compute the same value.  The
// 'from' end stores req[l].  Dynamic LDS: one real per item of the block.
// The blocks after the node blocks (kLinkThreads inlet cells each) apply the
// inlet cells' new depths.
template <typename real, bool GATE>
__global__ void __launch_bounds__(kLinkThreads) k_link_node(const Ctl* __restrict__ ctls, KP<real> p,
        DrainDev<real> d, int node_blocks, typename V4<real>::type* __restrict__ state) {
    constexpr int KB = kRowEntries;
    extern __shared__ __align__(16) unsigned char fh_lds_dyn[];
    const int bid = xcd_block(blockIdx.x, gridDim.x, p.remap);
    if (bid >= node_blocks) {
        // the inlet cells' depths of this step's inlet exchange (k_inlet leaves
        // the state unchanged for the sibling requests)
        const int j = (bid - node_blocks) * (int)blockDim.x + (int)threadIdx.x;
        if (j >= d.Ci) return;
        const int mi = d.seg_icell.blk_member[j / kInletRows];
        if (GATE && !ctls[mi].active) return;
        const int c = d.ic_cells[j];
        const real h = d.ic_h[j];
        FH_ISSUE_LOADS_OF("s"(state));
        if (!ctls[mi].active) return;
        if (j < d.seg_icell.mem_end[mi]) state[c].x = h;
        return;
    }
    real* rq = reinterpret_cast<real*>(fh_lds_dyn);
    const int b = bid;
    const int t = threadIdx.x;
    const int mem = d.seg_node.blk_member[b];
    if (GATE && !ctls[mem].active) return;      // the whole block: one member
    const int mend = d.seg_node.mem_end[mem];
    const int N = d.N;
    const int it0 = d.lb_ptr[b], it1 = d.lb_ptr[b + 1];
    // this thread's first item and its end nodes' data, loaded ahead (threads
    // past the block's items read item 0, never used)
    const int q = it0 + t < it1 ? it0 + t : 0;
    const int l = d.it_l[q], na = d.it_a[q], nb = d.it_b[q];
    const int a0 = d.it_a0[q], a1 = d.it_a1[q], b0 = d.it_b0[q], b1 = d.it_b1[q];
    const LinkData<real> lk = load_link(d, l);
    const real Va = d.V[na], inva = d.inv[na], Asa = d.As[na];
    const real Vb = d.V[nb], invb = d.inv[nb], Asb = d.As[nb];
    const Nets2<real> neta = first_nets(d, a0, a1), netb = first_nets(d, b0, b1);
    // this thread's node (threads < kNodeBlock): inlets, volume, first KB entries
    const bool node_thread = t < kNodeBlock;
    const int n = b * kNodeBlock + (node_thread ? t : 0);
    const int i0 = d.ni_ptr[n], i1 = d.ni_ptr[n + 1];
    const int k0 = d.nl_ptr[n], k1 = d.nl_ptr[n + 1];
    const real vol = d.V[n];
    const Nets2<real> net = first_nets(d, i0, i1);
    int en[KB], src[KB];
#pragma unroll
    for (int u = 0; u < KB; ++u) {
        en[u] = d.nle[u * N + n];
        src[u] = d.nls[u * N + n];
    }
    FH_ISSUE_LOADS();
    const Ctl* ctl = ctls + mem;
    if (!ctl->active) return;                 // the whole block: one member
    const real dt = (real)ctl->dt_real;
    // the node's own post-inlet volume ahead of the barrier
    const bool live_n = node_thread && n < mend;
    const real vn = live_n ? post_inlet_volume(d, i0, i1, vol, net.n0, net.n1) : real(0);
    if (it0 + t < it1) {
        const real va = post_inlet_volume(d, a0, a1, Va, neta.n0, neta.n1);
        const real vb = na == nb ? va : post_inlet_volume(d, b0, b1, Vb, netb.n0, netb.n1);
        rq[t] = link_request(p, lk, head_of(inva, Asa, va), head_of(invb, Asb, vb), dt);
    }
    for (int i = it0 + t + kLinkThreads; i < it1; i += kLinkThreads) {
        const int li = d.it_l[i], ai = d.it_a[i], bi = d.it_b[i];
        const int ia0 = d.it_a0[i], ia1 = d.it_a1[i], ib0 = d.it_b0[i], ib1 = d.it_b1[i];
        const Nets2<real> na2 = first_nets(d, ia0, ia1), nb2 = first_nets(d, ib0, ib1);
        const real va = post_inlet_volume(d, ia0, ia1, d.V[ai], na2.n0, na2.n1);
        const real vb = ai == bi ? va : post_inlet_volume(d, ib0, ib1, d.V[bi], nb2.n0, nb2.n1);
        rq[i - it0] = link_request(p, load_link(d, li), head_of(d.inv[ai], d.As[ai], va),
                                   head_of(d.inv[bi], d.As[bi], vb), dt);
    }
    __syncthreads();
    if (!live_n) return;
    d.Vp[n] = vn;
    real total = real(0);
    for (int kb = k0; kb < k1; kb += KB) {
        if (kb != k0) {
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                if (kb + u >= k1) break;
                en[u] = d.nl_ent[kb + u];
                src[u] = d.nl_src[kb + u];
            }
        }
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            if (kb + u >= k1) break;
            const bool is_to = en[u] & 1;
            const real r = rq[src[u]];
            if (!is_to) d.req[en[u] >> 1] = r;
            const bool positive = r >= real(0);
            if (positive != is_to) total = total + fabs(r);   // donor = from if positive else to
        }
    }
    d.sur_factor[n] = vmin(real(1), Divisor<real>(vmax(total, real(1.0e-12))).div(vn));
}

// Node continuity (one thread per node): donor-limited link transfers
// (reference scatter order: all 'from' links, then all 'to' links), outfall,
// overflow above storage capacity (weighted entry table, then the node surface
// cell fallback); writes the link flow for links this node is 'from' of.
// A node whose fallback volume is the only return its surface cell receives
// (nr_cell >= 0) applies it here, as k_return would: h + (0 + fb / A).
template <typename real, bool GATE>
__global__ void __launch_bounds__(256) k_node_update(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
                                                    typename V4<real>::type* __restrict__ state,
                                                    double* __restrict__ part) {
    const int b = xcd_block(blockIdx.x, gridDim.x, p.remap);
    const int n = b * blockDim.x + threadIdx.x;
    const int mem = d.seg_node.blk_member[b];
    if (GATE && !ctls[mem].active) return;
    const real zero = real(0);
    const int k0 = d.nl2_ptr[n], k1 = d.nl2_ptr[n + 1];
    const bool dry = d.dry != 0;
    const bool entries = d.S > 0 && !dry;
    const int e0 = entries ? d.ne_ptr[n] : 0, e1 = entries ? d.ne_ptr[n + 1] : 0;
    const real vp = d.Vp[n], f_self = d.sur_factor[n];
    const real inv = d.inv[n], As = d.As[n], dmax = d.dmax[n];
    const int flag = d.nflag[n];
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
    const real sw0 = entries ? d.nesw0[n] : zero;
    const int rcell = dry ? -1 : d.nr_cell[n];
    const real rA = d.nr_area[n];
    const real rh = state[rcell >= 0 ? rcell : 0].x;
    SlotAcc<3> acc3;
    acc3.prefetch(part, b);
    FH_ISSUE_LOADS_OF("s"(part), "s"(state));
    Ctl* ctl = ctls + mem;
    if (!ctl->active) return;
    double out_sum = 0.0, total_sum = 0.0, unret_sum = 0.0;
    if (n < d.seg_node.mem_end[mem]) {
        const real dt = (real)ctl->dt_real;
        real vol = vp;
        double moved = 0.0;    // sum of |transfer| over the links this node is 'from' of
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
                        d.Q[en[u] >> 1] = Divisor<real>(vmax(dt, real(1.0e-12))).div(T);
                        if (dry) moved += (double)fabs(T);
                    } else {
                        delta = delta + T;
                    }
                }
            }
            vol = vmax(vol + delta, zero);
        }
        // outfall (v1_parity.py:361-367); a node without an outfall slot
        // discharges exactly zero
        real outf = zero;
        if (flag >> 1) {
            const typename V4<real>::type o = d.outfall[(flag >> 1) - 1];
            const real hd = head_of(inv, As, vol) - (d.dry_tw ? d.dry_tw[n] : o.x);
            const real sgn = hd > zero ? real(1) : (hd < zero ? real(-1) : zero);
            real oq = o.y * o.z * sqrt(p.two_g * fabs(hd)) * sgn;
            oq = vmin(oq, vmax(o.w, zero));
            outf = vmin(vol, vmax(oq, zero) * dt);
            vol = vmax(vol - outf, zero);
        }
        const real capacity = As * vmax(dmax, zero);
        if (dry) {
            // dry weather: overflow reported, capped when asked, never returned
            const real over = vmax(vol - capacity, zero);
            if (d.dry_cap) vol = vmin(vol, capacity);
            d.V[n] = vol;
            out_sum = (double)outf;
            total_sum = (double)over;
            unret_sum = moved;
        } else {
        // overflow (v1_parity.py:371-404)
        const real over = vmax(vol - capacity, zero);
        real returned = zero;
        real wsum = zero;
        if (d.S > 0) {
            for (int k = e0; k < e1; ++k) wsum = wsum + vmax(k == e0 ? sw0 : d.ne_sw[k], zero);
            for (int k = e0; k < e1; ++k) {
                const real raw = vmax(k == e0 ? sw0 : d.ne_sw[k], zero);
                const real norm = d.snorm[mem] ? raw : Divisor<real>(vmax(wsum, real(1.0e-12))).div(raw);
                returned = returned + over * norm;
            }
            returned = vmin(returned, over);
        }
        vol = vol - returned;
        const real residual = vmax(vol - capacity, zero);
        const real fb = (flag & 1) ? residual : zero;
        vol = vol - fb;
        d.V[n] = vol;
        if (d.store_returns) {
            d.overflow[n] = over;
            d.wsum[n] = wsum;
            d.fallback[n] = fb;
        }
        // h + (0 + fb / A) leaves h's bits unchanged when fb is +0, unless h is -0
        if (rcell >= 0 && (fb != zero || (rh == zero && signbit(rh)))) {
            real acc = zero;
            acc = acc + Divisor<real>(rA).div(fb);
            state[rcell].x = rh + acc;
        }
        out_sum = (double)outf;
        total_sum = (double)(returned + fb);
        unret_sum = (double)vmax(vol - capacity, zero);
        }
    }
    __shared__ double lds[48];
    double v[3] = {out_sum, total_sum, unret_sum};
    block_sum<3>(v, lds);
    acc3.commit(part, b, v);
}

// Return pass (one thread per receiving cell of rx_rows): weighted entry
// returns, node fallback returns, then recharge capture/loss
// (v1_parity.py:371-404, 415-427).
template <typename real, bool GATE>
__global__ void __launch_bounds__(256) k_return(Ctl* __restrict__ ctls, KP<real> p, DrainDev<real> d,
        typename V4<real>::type* __restrict__ state, const real* __restrict__ area, double* __restrict__ part) {
    const int b = xcd_block(blockIdx.x, gridDim.x, p.remap);
    const int x = b * blockDim.x + threadIdx.x;
    const int mem = d.seg_rx.blk_member[b];
    if (GATE && !ctls[mem].active) return;
    const int j = d.rx_rows[x];
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
    if (x < d.seg_rx.mem_end[mem]) {
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
                const real norm = d.snorm[mem] ? raw : Divisor<real>(vmax(ws, real(1.0e-12))).div(raw);
                acc = acc + Divisor<real>(A).div(ov * norm);
            }
            h = h + acc;
        }
        {
            real acc = zero;
            for (int k = kf0; k < kf1; ++k) acc = acc + Divisor<real>(A).div(k == kf0 ? fb0 : d.fallback[d.rc_fidx[k]]);
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
            h = vmax(h - Divisor<real>(A).div(removed), zero);
        }
        state[c].x = h;
    }
    __shared__ double lds[32];
    double v[2] = {cap_sum, loss_sum};
    block_sum<2>(v, lds);
    acc2.commit(part, b, v);
}

// Dry-weather step prologue (triangular_backend.py:862-871): V = max(V, 0) +
// max(inflow, 0) dt for every node, every inlet slot's exchange zeroed (no
// surface exchange), and the member's step dt for the link and node kernels.
template <typename real>
__global__ void __launch_bounds__(256) k_dw_prep(Ctl* __restrict__ ctls, DrainDev<real> d, int n_slots,
                                                 const real* __restrict__ inflow, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const real dtr = (real)dt;
    if (i < d.N) {
        real v = vmax(d.V[i], real(0));
        if (inflow) v = v + vmax(inflow[i], real(0)) * dtr;
        d.V[i] = v;
    }
    if (i < n_slots) d.ni_net[i] = real(0);
    if (i == 0) {
        ctls[0].dt = (double)dtr;
        ctls[0].dt_real = (double)dtr;
        ctls[0].active = 1;
    }
}

}  // namespace fh
