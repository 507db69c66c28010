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
    int N, I, Ci, L, S, Cr, R, Rcv, Cx;
    // nodes
    const real *inv, *As, *dmax, *o_area, *o_coef, *o_tw, *o_cap;
    const int* nsc;
    real* V;
    real *sur_factor, *overflow, *fallback, *wsum;
    const int* explicit_;  // node has valid entry-table rows
    const int *ni_ptr, *ni_idx;        // node -> inlets
    const int *nl_ptr, *nl_ent;        // node -> incident links, ascending link, (l << 1) | is_to
    const int *ne_ptr, *ne_idx;        // node -> valid entries
    // inlets
    const int *icell, *inode;
    const real *capq, *surq;
    real *sur, *cap;
    const int *ic_cells, *ic_ptr, *ic_idx;  // inlet cells -> inlets
    // links
    const int *from, *to, *geom;
    const real *len, *man, *lcap, *linv, *D, *minor, *rw, *rh;
    real *Q, *req;
    // entries
    const int *snode;
    const real* sw;
    int snorm;
    // return cells
    const int *rc_cells, *rc_eptr, *rc_eidx, *rc_fptr, *rc_fidx, *rc_rptr, *rc_ridx;
    // recharge rows
    const real *rS, *rcc, *rrc, *rstop;
    real* rstore;
    // external
    const float* ext_q;      // [T][Rcv]
    const float* ext_q_override;
    int use_override;
    const int *xc_cells, *xc_ptr, *xc_idx;
};

template <typename real>
__device__ __forceinline__ real node_head(const DrainDev<real>& d, int n, real vol) {
    return d.inv[n] + vmax(vol, real(0)) / vmax(d.As[n], real(1.0e-12));
}

template <typename real>
__device__ __forceinline__ real orifice_q(const KP<real>& p, real dh) {
    const real sgn = dh > real(0) ? real(1) : (dh < real(0) ? real(-1) : real(0));
    return p.capture_q * sqrt(p.two_g * fabs(dh)) * sgn;
}

// commit helper used by the last block of each ledger kernel
template <int NT>
__device__ __forceinline__ void commit_terms(Ctl* ctl, int counter, const double* part, int nparts,
                                             const int (&terms)[NT], double* lds) {
    double s[NT];
    sum_partials<NT>(part, nparts, s, lds);
    if (threadIdx.x == 0) {
        ctl->counters[counter] = 0u;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            ctl->diag[terms[k]] = s[k];
            ctl->ledger[terms[k]] += s[k];
        }
    }
}

// v1_parity.py:407-412 -- h[recv] += max(q, 0) dt / A, q interpolated at t + dt/2.
template <typename real>
__global__ void __launch_bounds__(256) k_external(Ctl* __restrict__ ctl, KP<real> p, DrainDev<real> d,
                                                 typename V4<real>::type* __restrict__ state,
                                                 const real* __restrict__ area, double* __restrict__ part) {
    if (!ctl->active) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const real dt = (real)ctl->dt_real;
    double added_sum = 0.0;
    if (j < d.Cx) {
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
                const real ql = (real)d.ext_q[(size_t)lo * d.Rcv + r];
                const real qu = (real)d.ext_q[(size_t)up * d.Rcv + r];
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
    publish_partials<1>(part, blockIdx.x, v);
    if (!arrive_last(&ctl->counters[C_EXTERNAL], gridDim.x)) return;
    const int terms[1] = {L_EXTERNAL};
    commit_terms<1>(ctl, C_EXTERNAL, part, gridDim.x, terms, lds);
}

// Inlet exchange, pass A (one thread per node): surcharge requests of the node's
// inlets and the node's donor factor  min(1, V / sum(requests)).
template <typename real>
__global__ void __launch_bounds__(256) k_inlet_node_factor(const Ctl* __restrict__ ctl, KP<real> p,
        DrainDev<real> d, const typename V4<real>::type* __restrict__ state) {
    if (!ctl->active) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.N) return;
    const int k0 = d.ni_ptr[n], k1 = d.ni_ptr[n + 1];
    if (k0 == k1) return;
    const real dt = (real)ctl->dt_real;
    const real vol = d.V[n];
    const real nh = node_head(d, n, vol);
    real total = real(0);
    for (int k = k0; k < k1; ++k) {
        const int i = d.ni_idx[k];
        const int c = d.icell[i];
        const typename V4<real>::type s = state[c];
        const real q = orifice_q(p, (s.w + s.x) - nh);
        const real rs = vmin(vmax(-q, real(0)) * dt, vmax(d.surq[i], real(0)) * dt);
        d.sur[i] = rs;
        total = total + rs;
    }
    const real factor = vmin(real(1), vol / vmax(total, real(1.0e-12)));
    for (int k = k0; k < k1; ++k) {
        const int i = d.ni_idx[k];
        d.sur[i] = d.sur[i] * factor;
    }
}

// Inlet exchange, pass B (one thread per inlet cell): capture requests, the
// cell's donor factor min(1, h A / sum), captured volumes, and the depth update
// h = max(h + sum((sur - cap) / A), 0).
template <typename real>
__global__ void __launch_bounds__(256) k_inlet_cell(Ctl* __restrict__ ctl, KP<real> p, DrainDev<real> d,
        typename V4<real>::type* __restrict__ state, const real* __restrict__ area, double* __restrict__ part) {
    if (!ctl->active) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double cap_sum = 0.0, sur_sum = 0.0;
    if (j < d.Ci) {
        const int c = d.ic_cells[j];
        const real dt = (real)ctl->dt_real;
        typename V4<real>::type s = state[c];
        const real A = area[c];
        const real sh = s.w + s.x;
        const int k0 = d.ic_ptr[j], k1 = d.ic_ptr[j + 1];
        real total = real(0);
        for (int k = k0; k < k1; ++k) {
            const int i = d.ic_idx[k];
            const int n = d.inode[i];
            const real q = orifice_q(p, sh - node_head(d, n, d.V[n]));
            const real rc = vmin(vmax(q, real(0)) * dt, vmax(d.capq[i], real(0)) * dt);
            d.cap[i] = rc;
            total = total + rc;
        }
        const real factor = vmin(real(1), (s.x * A) / vmax(total, real(1.0e-12)));
        real acc = real(0);
        for (int k = k0; k < k1; ++k) {
            const int i = d.ic_idx[k];
            const real cp = d.cap[i] * factor;
            d.cap[i] = cp;
            const real sr = d.sur[i];
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
    publish_partials<2>(part, blockIdx.x, v);
    if (!arrive_last(&ctl->counters[C_INLET], gridDim.x)) return;
    const int terms[2] = {L_CAPTURE, L_SURCHARGE};
    commit_terms<2>(ctl, C_INLET, part, gridDim.x, terms, lds);
}

// Inlet exchange, pass C (one thread per node with inlets): V = max(V + sum(cap - sur), 0).
template <typename real>
__global__ void __launch_bounds__(256) k_inlet_node_apply(const Ctl* __restrict__ ctl, DrainDev<real> d) {
    if (!ctl->active) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.N) return;
    const int k0 = d.ni_ptr[n], k1 = d.ni_ptr[n + 1];
    if (k0 == k1) return;
    real acc = real(0);
    for (int k = k0; k < k1; ++k) {
        const int i = d.ni_idx[k];
        acc = acc + (d.cap[i] - d.sur[i]);
    }
    d.V[n] = vmax(d.V[n] + acc, real(0));
}

template <typename real>
__device__ __forceinline__ void circular_theta(real depth, real diameter, real& radius, real& theta) {
    radius = real(0.5) * diameter;
    const real od = vmin(vmax(depth, real(0)), diameter);
    real x = (radius - od) / radius;
    x = vmin(vmax(x, real(-1)), real(1));
    theta = real(2) * acos(x);
}

// Pipe link (one thread per link): semi-implicit Saint-Venant momentum with
// Manning friction and minor losses, clipped to capacity; stores the signed
// requested transfer Q dt.
template <typename real>
__global__ void __launch_bounds__(256) k_link(const Ctl* __restrict__ ctl, KP<real> p, DrainDev<real> d) {
    if (!ctl->active) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.L) return;
    const real dt = (real)ctl->dt_real;
    const real zero = real(0), g = p.g;
    const int a = d.from[l], b = d.to[l];
    const real ha = node_head(d, a, d.V[a]), hb = node_head(d, b, d.V[b]);
    const real depth = vmax(real(0.5) * (ha + hb) - d.linv[l], zero);
    const real Draw = d.D[l];
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
    if (d.geom[l] == 2) {
        const real od = vmin(vmax(depth, zero), vmax(d.rh[l], real(1.0e-6)));
        const real oa = vmax(d.rw[l], real(1.0e-6)) * od;
        flow_area = oa;
        radius = oa / vmax(d.rw[l] + real(2) * od, real(1.0e-12));
    }
    const real length = vmax(d.len[l], real(1.0e-6));
    const real area = vmax(flow_area, real(1.0e-12));
    radius = vmax(radius, real(1.0e-12));
    const real slope = (ha - hb) / length;
    const real nm = vmax(d.man[l], zero);
    const real q0 = d.Q[l];
    const real absq = fabs(q0);
    const real friction = g * (nm * nm) * absq / (pow(radius, p.pow43) * area);
    const real minor = vmax(d.minor[l], zero) * absq / (real(2) * length * area);
    real cand = (q0 + g * area * dt * slope) / (real(1) + dt * (friction + minor));
    const real cap = vmax(d.lcap[l], zero);
    cand = vmin(vmax(cand, -cap), cap);
    d.req[l] = cand * dt;
}

// Node donor factor (one thread per node): sum of |requested| over the links
// this node donates to (ascending link order), factor min(1, V / sum).
template <typename real>
__global__ void __launch_bounds__(256) k_node_factor(const Ctl* __restrict__ ctl, DrainDev<real> d) {
    if (!ctl->active) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.N) return;
    real total = real(0);
    for (int k = d.nl_ptr[n]; k < d.nl_ptr[n + 1]; ++k) {
        const int en = d.nl_ent[k];
        const int l = en >> 1;
        const bool is_to = en & 1;
        const real r = d.req[l];
        const bool positive = r >= real(0);
        if (positive != is_to) total = total + fabs(r);   // donor = from if positive else to
    }
    d.sur_factor[n] = vmin(real(1), d.V[n] / vmax(total, real(1.0e-12)));
}

// Node continuity (one thread per node): donor-limited link transfers
// (reference scatter order: all 'from' links, then all 'to' links), outfall,
// overflow above storage capacity (weighted entry table, then the node surface
// cell fallback); writes the link flow for links this node is 'from' of.
template <typename real>
__global__ void __launch_bounds__(256) k_node_update(Ctl* __restrict__ ctl, KP<real> p, DrainDev<real> d,
                                                    double* __restrict__ part) {
    if (!ctl->active) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const real zero = real(0);
    double out_sum = 0.0, total_sum = 0.0, unret_sum = 0.0;
    if (n < d.N) {
        const real dt = (real)ctl->dt_real;
        real vol = d.V[n];
        if (d.L > 0) {
            const int k0 = d.nl_ptr[n], k1 = d.nl_ptr[n + 1];
            real delta = zero;
            for (int pass = 0; pass < 2; ++pass) {
                for (int k = k0; k < k1; ++k) {
                    const int en = d.nl_ent[k];
                    const int is_to = en & 1;
                    if (is_to != pass) continue;
                    const int l = en >> 1;
                    const real r = d.req[l];
                    const bool positive = r >= zero;
                    const int donor = positive ? d.from[l] : d.to[l];
                    const real T = (positive ? real(1) : real(-1)) * fabs(r) * d.sur_factor[donor];
                    if (!is_to) {
                        delta = delta + (-T);
                        d.Q[l] = T / vmax(dt, real(1.0e-12));
                    } else {
                        delta = delta + T;
                    }
                }
            }
            vol = vmax(vol + delta, zero);
        }
        // outfall (v1_parity.py:361-367)
        const real hd = node_head(d, n, vol) - d.o_tw[n];
        const real sgn = hd > zero ? real(1) : (hd < zero ? real(-1) : zero);
        real oq = d.o_coef[n] * d.o_area[n] * sqrt(p.two_g * fabs(hd)) * sgn;
        oq = vmin(oq, vmax(d.o_cap[n], zero));
        const real outf = vmin(vol, vmax(oq, zero) * dt);
        vol = vmax(vol - outf, zero);
        // overflow (v1_parity.py:371-404)
        const real capacity = d.As[n] * vmax(d.dmax[n], zero);
        const real over = vmax(vol - capacity, zero);
        real returned = zero;
        real wsum = zero;
        if (d.S > 0) {
            const int k0 = d.ne_ptr[n], k1 = d.ne_ptr[n + 1];
            for (int k = k0; k < k1; ++k) wsum = wsum + vmax(d.sw[d.ne_idx[k]], zero);
            for (int k = k0; k < k1; ++k) {
                const real raw = vmax(d.sw[d.ne_idx[k]], zero);
                const real norm = d.snorm ? raw : raw / vmax(wsum, real(1.0e-12));
                returned = returned + over * norm;
            }
            returned = vmin(returned, over);
        }
        vol = vol - returned;
        const real residual = vmax(vol - capacity, zero);
        const real fb = (!d.explicit_[n] && d.nsc[n] >= 0) ? residual : zero;
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
    publish_partials<3>(part, blockIdx.x, v);
    if (!arrive_last(&ctl->counters[C_NODE], gridDim.x)) return;
    const int terms[3] = {L_OUTFALL, L_OVERFLOW, L_UNRETURNED};
    commit_terms<3>(ctl, C_NODE, part, gridDim.x, terms, lds);
}

// Return pass (one thread per receiving cell): weighted entry returns, node
// fallback returns, then recharge capture/loss (v1_parity.py:415-427).
template <typename real>
__global__ void __launch_bounds__(256) k_return(Ctl* __restrict__ ctl, KP<real> p, DrainDev<real> d,
        typename V4<real>::type* __restrict__ state, const real* __restrict__ area, double* __restrict__ part) {
    if (!ctl->active) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const real zero = real(0);
    double cap_sum = 0.0, loss_sum = 0.0;
    if (j < d.Cr) {
        const int c = d.rc_cells[j];
        const real dt = (real)ctl->dt_real;
        const real A = area[c];
        real h = state[c].x;
        if (d.S > 0) {
            real acc = zero;
            for (int k = d.rc_eptr[j]; k < d.rc_eptr[j + 1]; ++k) {
                const int s = d.rc_eidx[k];
                const int n = d.snode[s];
                const real raw = vmax(d.sw[s], zero);
                const real norm = d.snorm ? raw : raw / vmax(d.wsum[n], real(1.0e-12));
                acc = acc + (d.overflow[n] * norm) / A;
            }
            h = h + acc;
        }
        {
            real acc = zero;
            for (int k = d.rc_fptr[j]; k < d.rc_fptr[j + 1]; ++k) {
                const int n = d.rc_fidx[k];
                acc = acc + d.fallback[n] / A;
            }
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
    publish_partials<2>(part, blockIdx.x, v);
    if (!arrive_last(&ctl->counters[C_RETURN], gridDim.x)) return;
    const int terms[2] = {L_RCAPTURE, L_RLOSS};
    commit_terms<2>(ctl, C_RETURN, part, gridDim.x, terms, lds);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Counting-sort CSR: rows[i] in [0, nrows) for items i (negative rows skipped);
// items of a row stay in ascending item order.
inline void build_csr(const std::vector<int>& rows, int nrows, std::vector<int>& ptr, std::vector<int>& idx,
                      const std::vector<int>* payload = nullptr) {
    ptr.assign((size_t)nrows + 1, 0);
    for (int r : rows)
        if (r >= 0) ptr[(size_t)r + 1]++;
    for (int r = 0; r < nrows; ++r) ptr[(size_t)r + 1] += ptr[r];
    idx.assign((size_t)ptr[nrows], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < rows.size(); ++i)
        if (rows[i] >= 0) idx[(size_t)fill[rows[i]]++] = payload ? (*payload)[i] : (int)i;
}

template <typename real>
struct Drainage {
    template <typename T>
    struct Arr {
        T* p = nullptr;
        size_t n = 0;
        void set(const std::vector<T>& v, hipStream_t s) {
            if (v.size() != n) {
                if (p) (void)hipFree(p);
                p = nullptr;
                n = v.size();
                if (n && hipMalloc((void**)&p, n * sizeof(T)) != hipSuccess) throw std::bad_alloc();
            }
            if (n && hipMemcpyAsync(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
                throw std::runtime_error("hipMemcpyAsync failed");
        }
        void get(std::vector<T>& v, hipStream_t s) const {
            v.resize(n);
            if (n && hipMemcpyAsync(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess)
                throw std::runtime_error("hipMemcpyAsync failed");
        }
        ~Arr() { if (p) (void)hipFree(p); }
    };
    DrainDev<real> dd{};
    int64_t N = 0, I = 0, L = 0, S = 0, R = 0, Rcv = 0, T = 0;
    int Ci = 0, Cr = 0, Cx = 0;
    // host copies needed to rebuild the return-cell tables
    std::vector<int> h_nsc, h_scell, h_snode, h_explicit, h_rcell, h_rcv;
    std::vector<real> h_sw;
    Arr<real> inv, As, dmax, oarea, ocoef, otw, ocap, V, sfac, over, fb, wsum;
    Arr<int> nsc, expl, ni_ptr, ni_idx, nl_ptr, nl_ent, ne_ptr, ne_idx;
    Arr<int> icell, inode, ic_cells, ic_ptr, ic_idx;
    Arr<real> capq, surq, sur, cap;
    Arr<int> from, to, geom;
    Arr<real> len, man, lcap, linv, D, minor, rw, rh, Q, req;
    Arr<int> snode;
    Arr<real> sw;
    Arr<int> rc_cells, rc_eptr, rc_eidx, rc_fptr, rc_fidx, rc_rptr, rc_ridx;
    Arr<real> rS, rcc, rrc, rstop, rstore;
    Arr<float> ext_t, ext_q, ext_ov;
    Arr<int> xc_cells, xc_ptr, xc_idx;
    Arr<double> part_inlet, part_node, part_ret, part_ext;
    std::vector<float> ov_host;

    template <typename V_>
    static std::vector<real> conv(const float* src, int64_t n, float dflt = 0.0f) {
        std::vector<real> v((size_t)n);
        for (int64_t i = 0; i < n; ++i) v[(size_t)i] = (real)(src ? src[i] : dflt);
        return v;
    }
    static std::vector<int> ints(const int32_t* src, int64_t n, int dflt = -1) {
        std::vector<int> v((size_t)n);
        for (int64_t i = 0; i < n; ++i) v[(size_t)i] = src ? src[i] : dflt;
        return v;
    }

    template <class Eng>
    void upload_coupling(Eng& eng, int64_t N_, int64_t I_, int64_t S_, const int32_t* icell_,
                         const int32_t* inode_, const int32_t* nsc_, const float* inv_, const float* As_,
                         const float* dmax_, const float* capq_, const float* surq_, const int32_t* scell_,
                         const int32_t* snode_, const float* sw_, int32_t snorm_) {
        hipStream_t s = eng.stream;
        if (N_ < 0 || I_ < 0 || S_ < 0) throw ValueError{"negative coupling sizes"};
        if (N_ > 0 && (!inv_ || !As_ || !dmax_)) throw ValueError{"node arrays must not be NULL"};
        if (I_ > 0 && (!icell_ || !inode_)) throw ValueError{"inlet arrays must not be NULL"};
        if (S_ > 0 && (!scell_ || !snode_ || !sw_)) throw ValueError{"surcharge table arrays must not be NULL"};
        const int64_t C = eng.C;
        for (int64_t i = 0; i < I_; ++i)
            if (icell_[i] < 0 || icell_[i] >= C || inode_[i] < 0 || inode_[i] >= N_)
                throw ValueError{"inlet cell/node index out of range"};
        for (int64_t n = 0; n < N_; ++n)
            if (nsc_ && nsc_[n] >= C) throw ValueError{"node_surface_cell out of range"};
        for (int64_t k = 0; k < S_; ++k)
            if (scell_[k] >= C || snode_[k] >= N_) throw ValueError{"surcharge table index out of range"};
        N = N_; I = I_; S = S_;
        inv.set(conv<real>(inv_, N), s);
        As.set(conv<real>(As_, N), s);
        dmax.set(conv<real>(dmax_, N), s);
        h_nsc = ints(nsc_, N);
        nsc.set(h_nsc, s);
        if ((int64_t)V.n != N) {
            V.set(std::vector<real>((size_t)N, real(0)), s);
        }
        std::vector<real> zN((size_t)N, real(0));
        sfac.set(zN, s); over.set(zN, s); fb.set(zN, s); wsum.set(zN, s);
        if ((int64_t)oarea.n != N) {  // outfalls default to none until fh_upload_pipe
            oarea.set(zN, s); ocoef.set(zN, s); otw.set(zN, s); ocap.set(zN, s);
        }
        // inlets
        std::vector<int> ic = ints(icell_, I), in = ints(inode_, I);
        icell.set(ic, s); inode.set(in, s);
        capq.set(conv<real>(capq_, I), s);
        surq.set(conv<real>(surq_, I), s);
        std::vector<real> zI((size_t)std::max<int64_t>(I, 1), real(0));
        sur.set(zI, s); cap.set(zI, s);
        std::vector<int> p, x;
        build_csr(in, (int)N, p, x);
        ni_ptr.set(p, s); ni_idx.set(x.empty() ? std::vector<int>{0} : x, s);
        // inlet cells (distinct cells with inlets, ascending), CSR cell -> inlets ascending
        std::vector<int> cell_slot((size_t)C, -1), cells;
        for (int64_t i = 0; i < I; ++i)
            if (cell_slot[ic[i]] < 0) { cell_slot[ic[i]] = 0; }
        for (int64_t c = 0; c < C; ++c)
            if (cell_slot[c] >= 0) { cell_slot[c] = (int)cells.size(); cells.push_back((int)c); }
        Ci = (int)cells.size();
        std::vector<int> rows((size_t)I);
        for (int64_t i = 0; i < I; ++i) rows[i] = cell_slot[ic[i]];
        build_csr(rows, Ci, p, x);
        ic_cells.set(cells.empty() ? std::vector<int>{0} : cells, s);
        ic_ptr.set(p, s); ic_idx.set(x.empty() ? std::vector<int>{0} : x, s);
        // surcharge entry table
        h_scell = ints(scell_, S); h_snode = ints(snode_, S);
        h_sw = conv<real>(sw_, S);
        std::vector<int> valid_node((size_t)S);
        for (int64_t k = 0; k < S; ++k) valid_node[k] = (h_snode[k] >= 0 && h_scell[k] >= 0) ? h_snode[k] : -1;
        build_csr(valid_node, (int)N, p, x);
        ne_ptr.set(p, s); ne_idx.set(x.empty() ? std::vector<int>{0} : x, s);
        h_explicit.assign((size_t)N, 0);
        for (int64_t n = 0; n < N; ++n) h_explicit[n] = p[n + 1] > p[n] ? 1 : 0;
        expl.set(h_explicit.empty() ? std::vector<int>{0} : h_explicit, s);
        snode.set(h_snode.empty() ? std::vector<int>{0} : h_snode, s);
        sw.set(h_sw.empty() ? std::vector<real>{real(0)} : h_sw, s);
        dd.snorm = snorm_ ? 1 : 0;
        part_inlet.set(std::vector<double>(2 * (size_t)std::max(1, (Ci + 255) / 256), 0.0), s);
        part_node.set(std::vector<double>(3 * (size_t)std::max<int64_t>(1, (N + 255) / 256), 0.0), s);
        if (L == 0) {
            std::vector<int> pz((size_t)N + 1, 0);
            nl_ptr.set(pz, s); nl_ent.set(std::vector<int>{0}, s);
        }
        rebuild_return_cells(eng);
        refresh();
    }

    template <class Eng>
    void upload_pipe(Eng& eng, int64_t L_, const int32_t* from_, const int32_t* to_, const float* len_,
                     const float* n_, const float* cap_, const float* inv_, const float* D_, const float* minor_,
                     const int32_t* geom_, const float* rw_, const float* rh_, const float* oa, const float* oc,
                     const float* otw_, const float* ocap_) {
        hipStream_t s = eng.stream;
        if (L_ < 0) throw ValueError{"negative link count"};
        if (L_ > 0 && (!from_ || !to_ || !len_ || !n_ || !cap_ || !inv_ || !D_))
            throw ValueError{"link arrays must not be NULL"};
        for (int64_t l = 0; l < L_; ++l)
            if (from_[l] < 0 || from_[l] >= N || to_[l] < 0 || to_[l] >= N)
                throw ValueError{"link node index out of range (upload coupling first)"};
        L = L_;
        std::vector<int> f = ints(from_, L), t = ints(to_, L);
        from.set(f.empty() ? std::vector<int>{0} : f, s);
        to.set(t.empty() ? std::vector<int>{0} : t, s);
        std::vector<int> g = ints(geom_, L, 1);
        geom.set(g.empty() ? std::vector<int>{0} : g, s);
        auto setr = [&](Arr<real>& a, const float* src, float dflt = 0.0f) {
            auto v = conv<real>(src, L, dflt);
            if (v.empty()) v.push_back(real(0));
            a.set(v, s);
        };
        setr(len, len_); setr(man, n_); setr(lcap, cap_); setr(linv, inv_); setr(D, D_);
        setr(minor, minor_); setr(rw, rw_); setr(rh, rh_);
        std::vector<real> zL((size_t)std::max<int64_t>(L, 1), real(0));
        Q.set(zL, s); req.set(zL, s);
        // node -> incident links, ascending link index, orientation bit
        std::vector<int> cnt((size_t)N + 1, 0);
        for (int64_t l = 0; l < L; ++l) { cnt[f[l] + 1]++; cnt[t[l] + 1]++; }
        for (int64_t n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
        std::vector<int> ent((size_t)cnt[N]), fill(cnt.begin(), cnt.end() - 1);
        for (int64_t l = 0; l < L; ++l) {
            ent[fill[f[l]]++] = (int)(l << 1);
            ent[fill[t[l]]++] = (int)((l << 1) | 1);
        }
        nl_ptr.set(cnt, s);
        nl_ent.set(ent.empty() ? std::vector<int>{0} : ent, s);
        std::vector<real> vo = conv<real>(oa, N), vc = conv<real>(oc, N), vt = conv<real>(otw_, N),
                          vp = conv<real>(ocap_, N);
        oarea.set(vo, s); ocoef.set(vc, s); otw.set(vt, s); ocap.set(vp, s);
        refresh();
    }

    template <class Eng>
    void upload_recharge(Eng& eng, int64_t R_, const int32_t* cell, const float* S_, const float* cc,
                         const float* rc, const float* stop) {
        hipStream_t s = eng.stream;
        if (R_ < 0) throw ValueError{"negative recharge row count"};
        if (R_ > 0 && (!cell || !S_ || !cc || !rc || !stop)) throw ValueError{"recharge arrays must not be NULL"};
        for (int64_t r = 0; r < R_; ++r)
            if (cell[r] < 0 || cell[r] >= eng.C) throw ValueError{"recharge source_cell out of range"};
        R = R_;
        h_rcell = ints(cell, R);
        auto setr = [&](Arr<real>& a, const float* src) {
            auto v = conv<real>(src, R);
            if (v.empty()) v.push_back(real(0));
            a.set(v, s);
        };
        setr(rS, S_); setr(rcc, cc); setr(rrc, rc); setr(rstop, stop);
        rstore.set(std::vector<real>((size_t)std::max<int64_t>(R, 1), real(0)), s);
        rebuild_return_cells(eng);
        refresh();
    }

    // cells that receive entry returns, fallback returns or hold recharge rows
    template <class Eng>
    void rebuild_return_cells(Eng& eng) {
        hipStream_t s = eng.stream;
        const int64_t C = eng.C;
        std::vector<char> mark((size_t)C, 0);
        for (int64_t k = 0; k < S; ++k)
            if (h_snode[k] >= 0 && h_scell[k] >= 0) mark[h_scell[k]] = 1;
        for (int64_t n = 0; n < N; ++n)
            if (!h_explicit[n] && h_nsc[n] >= 0) mark[h_nsc[n]] = 1;
        for (int64_t r = 0; r < R; ++r) mark[h_rcell[r]] = 1;
        std::vector<int> slot((size_t)C, -1), cells;
        for (int64_t c = 0; c < C; ++c)
            if (mark[c]) { slot[c] = (int)cells.size(); cells.push_back((int)c); }
        Cr = (int)cells.size();
        std::vector<int> rows, p, x;
        rows.assign((size_t)S, -1);
        for (int64_t k = 0; k < S; ++k)
            if (h_snode[k] >= 0 && h_scell[k] >= 0) rows[k] = slot[h_scell[k]];
        build_csr(rows, Cr, p, x);
        rc_eptr.set(p, s); rc_eidx.set(x.empty() ? std::vector<int>{0} : x, s);
        rows.assign((size_t)N, -1);
        for (int64_t n = 0; n < N; ++n)
            if (!h_explicit[n] && h_nsc[n] >= 0) rows[n] = slot[h_nsc[n]];
        build_csr(rows, Cr, p, x);
        rc_fptr.set(p, s); rc_fidx.set(x.empty() ? std::vector<int>{0} : x, s);
        rows.assign((size_t)R, -1);
        for (int64_t r = 0; r < R; ++r) rows[r] = slot[h_rcell[r]];
        build_csr(rows, Cr, p, x);
        rc_rptr.set(p, s); rc_ridx.set(x.empty() ? std::vector<int>{0} : x, s);
        rc_cells.set(cells.empty() ? std::vector<int>{0} : cells, s);
        part_ret.set(std::vector<double>(2 * (size_t)std::max(1, (Cr + 255) / 256), 0.0), s);
    }

    template <class Eng>
    void set_external(Eng& eng, int64_t nt, const float* et, int64_t nrcv, const int32_t* rcv, const float* q) {
        hipStream_t s = eng.stream;
        if (nrcv > 0 && (nt <= 0 || !et || !rcv || !q)) throw ValueError{"external inflow arrays incomplete"};
        for (int64_t j = 0; j < nrcv; ++j)
            if (rcv[j] < 0 || rcv[j] >= eng.C) throw ValueError{"external receiver cell out of range"};
        Rcv = nrcv;
        T = nrcv > 0 ? nt : 0;
        std::vector<float> t(et ? et : nullptr, et ? et + T : nullptr);
        if (t.empty()) t.push_back(0.0f);
        ext_t.set(t, s);
        std::vector<float> qq;
        if (Rcv > 0) qq.assign(q, q + T * Rcv);
        if (qq.empty()) qq.push_back(0.0f);
        ext_q.set(qq, s);
        ext_ov.set(std::vector<float>((size_t)std::max<int64_t>(Rcv, 1), 0.0f), s);
        h_rcv = ints(rcv, Rcv);
        std::vector<int> slot((size_t)eng.C, -1), cells;
        std::vector<char> mark((size_t)eng.C, 0);
        for (int64_t j = 0; j < Rcv; ++j) mark[h_rcv[j]] = 1;
        for (int64_t c = 0; c < eng.C; ++c)
            if (mark[c]) { slot[c] = (int)cells.size(); cells.push_back((int)c); }
        Cx = (int)cells.size();
        std::vector<int> rows((size_t)Rcv), p, x;
        for (int64_t j = 0; j < Rcv; ++j) rows[j] = slot[h_rcv[j]];
        build_csr(rows, Cx, p, x);
        xc_cells.set(cells.empty() ? std::vector<int>{0} : cells, s);
        xc_ptr.set(p, s); xc_idx.set(x.empty() ? std::vector<int>{0} : x, s);
        part_ext.set(std::vector<double>((size_t)std::max(1, (Cx + 255) / 256), 0.0), s);
        refresh();
    }

    template <class Eng>
    void set_external_override(Eng& eng, const float* q) {
        if (Rcv == 0) return;
        ov_host.assign((size_t)Rcv, 0.0f);
        if (q)
            for (int64_t j = 0; j < Rcv; ++j) ov_host[j] = q[j];
        ext_ov.set(ov_host, eng.stream);
        dd.use_override = 1;
        eng.drop_graph();
    }
    template <class Eng>
    void clear_external_override(Eng& eng) {
        if (dd.use_override) {
            dd.use_override = 0;
            eng.drop_graph();
        }
    }

    template <class Eng>
    void set_state(Eng& eng, const float* vol, const float* q, const float* rs) {
        hipStream_t s = eng.stream;
        if (vol && N > 0) V.set(conv<real>(vol, N), s);
        if (q && L > 0) Q.set(conv<real>(q, L), s);
        if (rs && R > 0) rstore.set(conv<real>(rs, R), s);
        refresh();
    }

    template <class Eng>
    void get_state(Eng& eng, float* vol, float* q, float* rs) {
        hipStream_t s = eng.stream;
        std::vector<real> v;
        if (vol && N > 0) {
            V.get(v, s);
            if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("sync failed");
            for (int64_t n = 0; n < N; ++n) vol[n] = (float)v[n];
        }
        if (q && L > 0) {
            Q.get(v, s);
            if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("sync failed");
            for (int64_t l = 0; l < L; ++l) q[l] = (float)v[l];
        }
        if (rs && R > 0) {
            rstore.get(v, s);
            if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("sync failed");
            for (int64_t r = 0; r < R; ++r) rs[r] = (float)v[r];
        }
    }

    void refresh() {
        dd.N = (int)N; dd.I = (int)I; dd.Ci = Ci; dd.L = (int)L; dd.S = (int)S; dd.Cr = Cr; dd.R = (int)R;
        dd.Rcv = (int)Rcv; dd.Cx = Cx;
        dd.inv = inv.p; dd.As = As.p; dd.dmax = dmax.p; dd.o_area = oarea.p; dd.o_coef = ocoef.p;
        dd.o_tw = otw.p; dd.o_cap = ocap.p; dd.nsc = nsc.p; dd.V = V.p; dd.sur_factor = sfac.p;
        dd.overflow = over.p; dd.fallback = fb.p; dd.wsum = wsum.p; dd.explicit_ = expl.p;
        dd.ni_ptr = ni_ptr.p; dd.ni_idx = ni_idx.p; dd.nl_ptr = nl_ptr.p; dd.nl_ent = nl_ent.p;
        dd.ne_ptr = ne_ptr.p; dd.ne_idx = ne_idx.p;
        dd.icell = icell.p; dd.inode = inode.p; dd.capq = capq.p; dd.surq = surq.p; dd.sur = sur.p; dd.cap = cap.p;
        dd.ic_cells = ic_cells.p; dd.ic_ptr = ic_ptr.p; dd.ic_idx = ic_idx.p;
        dd.from = from.p; dd.to = to.p; dd.geom = geom.p; dd.len = len.p; dd.man = man.p; dd.lcap = lcap.p;
        dd.linv = linv.p; dd.D = D.p; dd.minor = minor.p; dd.rw = rw.p; dd.rh = rh.p; dd.Q = Q.p; dd.req = req.p;
        dd.snode = snode.p; dd.sw = sw.p;
        dd.rc_cells = rc_cells.p; dd.rc_eptr = rc_eptr.p; dd.rc_eidx = rc_eidx.p; dd.rc_fptr = rc_fptr.p;
        dd.rc_fidx = rc_fidx.p; dd.rc_rptr = rc_rptr.p; dd.rc_ridx = rc_ridx.p;
        dd.rS = rS.p; dd.rcc = rcc.p; dd.rrc = rrc.p; dd.rstop = rstop.p; dd.rstore = rstore.p;
        dd.ext_q = ext_q.p; dd.ext_q_override = ext_ov.p;
        dd.xc_cells = xc_cells.p; dd.xc_ptr = xc_ptr.p; dd.xc_idx = xc_idx.p;
    }

    const float* ext_times_ptr() const { return ext_t.p; }
    const real* volume_ptr() const { return V.p; }

    int64_t device_bytes() const {
        int64_t b = 0;
        b += (inv.n + As.n + dmax.n + oarea.n + ocoef.n + otw.n + ocap.n + V.n + sfac.n + over.n + fb.n + wsum.n) * sizeof(real);
        b += (nsc.n + expl.n + ni_ptr.n + ni_idx.n + nl_ptr.n + nl_ent.n + ne_ptr.n + ne_idx.n) * 4;
        b += (icell.n + inode.n + ic_cells.n + ic_ptr.n + ic_idx.n) * 4 + (capq.n + surq.n + sur.n + cap.n) * sizeof(real);
        b += (from.n + to.n + geom.n) * 4 + (len.n + man.n + lcap.n + linv.n + D.n + minor.n + rw.n + rh.n + Q.n + req.n) * sizeof(real);
        return b;
    }

    static int grid(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

    template <class Eng>
    void launch(Eng& eng) {
        hipStream_t s = eng.stream;
        if (Cx > 0)
            k_external<real><<<grid(Cx), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_ext.p);
        if (N == 0) return;
        if (Ci > 0) {
            k_inlet_node_factor<real><<<grid(N), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p);
            k_inlet_cell<real><<<grid(Ci), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_inlet.p);
            k_inlet_node_apply<real><<<grid(N), 256, 0, s>>>(eng.ctl, dd);
        }
        if (L > 0) {
            k_link<real><<<grid(L), 256, 0, s>>>(eng.ctl, eng.kp, dd);
            k_node_factor<real><<<grid(N), 256, 0, s>>>(eng.ctl, dd);
        }
        k_node_update<real><<<grid(N), 256, 0, s>>>(eng.ctl, eng.kp, dd, part_node.p);
        if (Cr > 0)
            k_return<real><<<grid(Cr), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_ret.p);
    }

    template <class Eng, class Timed>
    void profile(Eng& eng, Timed& timed) {
        hipStream_t s = eng.stream;
        const double rb = sizeof(real);
        if (Cx > 0)
            timed("external", (double)Rcv * 12.0 + Cx * 16.0, [&] {
                k_external<real><<<grid(Cx), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_ext.p);
            });
        if (N == 0) return;
        if (Ci > 0) {
            timed("inlet_node_factor", (double)N * (8 + 3 * rb) + (double)I * (12 + 3 * rb + 2 * rb), [&] {
                k_inlet_node_factor<real><<<grid(N), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p);
            });
            timed("inlet_cell", (double)Ci * (12 + 4 * rb) + (double)I * (8 + 5 * rb), [&] {
                k_inlet_cell<real><<<grid(Ci), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_inlet.p);
            });
            timed("inlet_node_apply", (double)N * (8 + 2 * rb) + (double)I * (4 + 2 * rb), [&] {
                k_inlet_node_apply<real><<<grid(N), 256, 0, s>>>(eng.ctl, dd);
            });
        }
        if (L > 0) {
            timed("link", (double)L * (12 + 11 * rb) + (double)N * 3 * rb, [&] {
                k_link<real><<<grid(L), 256, 0, s>>>(eng.ctl, eng.kp, dd);
            });
            timed("node_factor", (double)N * (4 + 2 * rb) + (double)L * (8 + rb), [&] {
                k_node_factor<real><<<grid(N), 256, 0, s>>>(eng.ctl, dd);
            });
        }
        timed("node_update", (double)N * (16 + 14 * rb) + (double)L * (16 + 3 * rb), [&] {
            k_node_update<real><<<grid(N), 256, 0, s>>>(eng.ctl, eng.kp, dd, part_node.p);
        });
        if (Cr > 0)
            timed("return", (double)Cr * (16 + 3 * rb) + (double)N * 2 * rb, [&] {
                k_return<real><<<grid(Cr), 256, 0, s>>>(eng.ctl, eng.kp, dd, eng.d_state.p, eng.d_area.p, part_ret.p);
            });
    }
};

}  // namespace fh
