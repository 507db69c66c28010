// floodhip: host engine and C ABI (see include/floodhip.h).
//
// One Engine<real> per context holds the device-resident scenario: SoA edge
// geometry, packed cell state (h, hu, hv, bed), the cell->edge gather table,
// drainage CSR tables, the step control block and the FP64 ledger.  A step is
// a fixed sequence of kernel launches on the context stream; fh_advance_to
// replays captured hipGraphs of steps_per_launch steps and reads the clock
// back once per replay batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/floodhip.h"
#include "fh_device.h"
#include "fh_surface.h"
#include "fh_drainage.h"

namespace {

thread_local std::string g_error;

struct HipError {
    hipError_t code;
    std::string where;
};

#define FH_CHECK(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) throw HipError{_e, std::string(#expr) + " @" + std::to_string(__LINE__)}; \
    } while (0)

using fh::ValueError;

template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        n = count;
        if (count) FH_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* host, size_t count, hipStream_t s) {
        if (count != n) alloc(count);
        if (count) FH_CHECK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T* host, size_t count, hipStream_t s) const {
        if (count) FH_CHECK(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    size_t bytes() const { return n * sizeof(T); }
    ~DevArray() { release(); }
};

struct KernelTiming {
    std::string name;
    double total_ms = 0.0;
    double bytes = 0.0;
    int64_t launches = 0;
};

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual void upload_surface(int64_t C, int64_t E, const float* bed, const float* n, const float* area,
                                const float* lc, const int32_t* el, const int32_t* er, const float* nx,
                                const float* ny, const float* len, const int32_t* code,
                                const uint8_t* active) = 0;
    virtual void upload_coupling(int64_t N, int64_t I, int64_t S, const int32_t* icell, const int32_t* inode,
                                 const int32_t* nsc, const float* inv, const float* As, const float* dmax,
                                 const float* capq, const float* surq, const int32_t* scell,
                                 const int32_t* snode, const float* sw, int32_t snorm) = 0;
    virtual void upload_pipe(int64_t L, const int32_t* from, const int32_t* to, const float* len,
                             const float* n, const float* cap, const float* inv, const float* D,
                             const float* minor, const int32_t* geom, const float* rw, const float* rh,
                             const float* oa, const float* oc, const float* otw, const float* ocap) = 0;
    virtual void upload_recharge(int64_t R, const int32_t* cell, const float* S, const float* cc,
                                 const float* rc, const float* stop) = 0;
    virtual void set_forcing(int64_t nr, const float* rs, const float* re, const float* rr, const float* m,
                             const float* f0, const float* fc, const float* k, int64_t nt, const float* et,
                             int64_t nrcv, const int32_t* rcv, const float* q) = 0;
    virtual void set_state(const float* h, const float* hu, const float* hv, const float* wet,
                           const float* vol, const float* q, const float* rs) = 0;
    virtual void get_state(float* h, float* hu, float* hv, float* wet, float* vol, float* q, float* rs) = 0;
    virtual void get_state_f64(double* h, double* hu, double* hv) = 0;
    virtual void set_clock(double t, double cached, int64_t step) = 0;
    virtual void get_clock(fh_clock* c) = 0;
    virtual void set_ledger(const double* l) = 0;
    virtual void get_ledger(double* l) = 0;
    virtual int advance_to(double target, double final_t, int64_t max_steps) = 0;
    virtual void step_once(double dt, double rain, const float* q, double* diag) = 0;
    virtual double stable_timestep() = 0;
    virtual std::vector<KernelTiming> profile(int steps) = 0;
    virtual int64_t device_bytes() const = 0;
};

template <typename real>
class Engine final : public EngineBase {
    typedef typename fh::V4<real>::type V;

public:
    explicit Engine(const fh_params& prm) : P(prm) {
        FH_CHECK(hipSetDevice(P.device));
        FH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        FH_CHECK(hipMalloc((void**)&ctl, sizeof(fh::Ctl)));
        FH_CHECK(hipHostMalloc((void**)&ctl_h, sizeof(fh::Ctl), hipHostMallocDefault));
        std::memset(ctl_h, 0, sizeof(fh::Ctl));
        ctl_h->cached_cfl = 1.0;
        ctl_h->eval_next = P.fixed_dt > 0.0 ? 0 : 1;
        FH_CHECK(hipMemcpyAsync(ctl, ctl_h, sizeof(fh::Ctl), hipMemcpyHostToDevice, stream));
        bs = std::max(64, std::min(1024, P.block_size > 0 ? P.block_size : 256));
        bs = (bs / 64) * 64;
        kp.g = (real)P.gravity;
        kp.eps = (real)P.min_depth;
        kp.half_g = (real)(0.5 * P.gravity);
        kp.two_g = (real)(2.0 * P.gravity);
        kp.cfl = (real)P.cfl;
        kp.max_speed = (real)P.max_speed;
        kp.tiny = (real)1.0e-12;
        kp.capture_q = (real)P.inlet_orifice_q;
        kp.pow43 = (real)(4.0 / 3.0);
        kp.reuse = (real)P.reuse_factor;
        kp.slot_coef = (real)(P.slot_wave_speed * P.slot_wave_speed);
        kp.max_dt = P.max_dt;
        kp.dt_floor = P.dt_floor;
        kp.fixed_dt = P.fixed_dt > 0.0 ? (double)(real)P.fixed_dt : 0.0;
        kp.interval = std::max(1, P.cfl_interval);
        kp.remap = P.xcd_remap;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    ~Engine() override {
        drop_graph();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ctl) (void)hipFree(ctl);
        if (ctl_h) (void)hipHostFree(ctl_h);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ------------------------------------------------------------------ uploads
    void upload_surface(int64_t C_, int64_t E_, const float* bed, const float* n, const float* area,
                        const float* lc_, const int32_t* el, const int32_t* er, const float* nx,
                        const float* ny, const float* len, const int32_t* code,
                        const uint8_t* active) override {
        if (C_ <= 0 || E_ < 0 || C_ > INT32_MAX / 2 || E_ > INT32_MAX / 2) throw ValueError{"surface sizes out of range"};
        if (!bed || !n || !area || !lc_ || (E_ && (!el || !er || !nx || !ny || !len || !code)))
            throw ValueError{"surface arrays must not be NULL"};
        drop_graph();
        C = C_;
        auto act = [&](int64_t c) { return active ? active[c] != 0 : true; };
        for (int64_t e = 0; e < E_; ++e) {
            if (el[e] < 0 || el[e] >= C || er[e] >= C) throw ValueError{"edge cell index out of range"};
        }
        // classify and keep edges whose left cell is active
        std::vector<int64_t> interior, ghost;
        std::vector<int32_t> ghost_kind;
        interior.reserve(E_);
        for (int64_t e = 0; e < E_; ++e) {
            const int32_t l = el[e], r = er[e];
            if (!act(l)) continue;
            const bool in = r >= 0;
            const bool ra = in && act(r);
            if (in && ra) {
                interior.push_back(e);
                continue;
            }
            const int32_t cd = (in && !ra) ? 0 : code[e];
            int32_t kind;
            if (cd == 0 || cd == 2) kind = fh::kGhostReflect;
            else if (!in && cd == 4) kind = fh::kGhostOutflow;
            else kind = fh::kGhostCopy;
            ghost.push_back(e);
            ghost_kind.push_back(kind);
        }
        E_int = (int64_t)interior.size();
        E = E_int + (int64_t)ghost.size();
        std::vector<int32_t> pos_of(E_, -1);
        std::vector<int32_t> hl(E), hr(E);
        std::vector<real> hnx(E), hny(E), hlen(E);
        for (int64_t i = 0; i < E_int; ++i) {
            int64_t e = interior[i];
            pos_of[e] = (int32_t)i;
            hl[i] = el[e]; hr[i] = er[e];
            hnx[i] = (real)nx[e]; hny[i] = (real)ny[e]; hlen[i] = (real)len[e];
        }
        for (size_t j = 0; j < ghost.size(); ++j) {
            int64_t e = ghost[j], i = E_int + (int64_t)j;
            pos_of[e] = (int32_t)i;
            hl[i] = el[e]; hr[i] = ghost_kind[j];
            hnx[i] = (real)nx[e]; hny[i] = (real)ny[e]; hlen[i] = (real)len[e];
        }
        // cell -> edge table: left-side entries (ascending original edge), then right-side
        std::vector<int32_t> lent(C * 3, -1), rent(C * 3, -1);
        std::vector<uint8_t> nl(C, 0), nr(C, 0);
        for (int64_t e = 0; e < E_; ++e) {
            const int32_t p = pos_of[e];
            if (p < 0) continue;
            const int32_t l = el[e];
            if (nl[l] + nr[l] >= 3) throw ValueError{"a cell has more than three edges"};
            lent[l * 3 + nl[l]++] = p << 1;
            if (p < E_int) {
                const int32_t r = er[e];
                if (nl[r] + nr[r] >= 3) throw ValueError{"a cell has more than three edges"};
                rent[r * 3 + nr[r]++] = (p << 1) | 1;
            }
        }
        std::vector<int32_t> c0(C), c1(C), c2(C);
        for (int64_t c = 0; c < C; ++c) {
            int32_t t[3] = {-1, -1, -1};
            int k = 0;
            for (int i = 0; i < nl[c]; ++i) t[k++] = lent[c * 3 + i];
            for (int i = 0; i < nr[c]; ++i) t[k++] = rent[c * 3 + i];
            if (!act(c)) t[0] = fh::kCellInactive;
            c0[c] = t[0]; c1[c] = t[1]; c2[c] = t[2];
        }
        std::vector<real> harea(C), hn2(C), hlc(C);
        host_bed.assign(C, real(0));
        for (int64_t c = 0; c < C; ++c) {
            const real nn = (real)n[c];
            harea[c] = (real)area[c];
            hn2[c] = nn * nn;
            hlc[c] = (real)lc_[c];
            host_bed[c] = (real)bed[c];
        }
        host_active.assign(C, 1);
        for (int64_t c = 0; c < C; ++c) host_active[c] = act(c) ? 1 : 0;
        d_el.upload(hl, stream); d_er.upload(hr, stream);
        d_nx.upload(hnx, stream); d_ny.upload(hny, stream); d_len.upload(hlen, stream);
        d_ce0.upload(c0, stream); d_ce1.upload(c1, stream); d_ce2.upload(c2, stream);
        d_area.upload(harea, stream); d_n2.upload(hn2, stream); d_lc.upload(hlc, stream);
        d_rec.alloc(std::max<int64_t>(E, 1));
        d_recR.alloc(std::max<int64_t>(E_int, 1));
        d_state.alloc(C);
        d_wet.alloc(C);
        if (P.compensated) d_comp.alloc(C);
        // zero state, bed in .w
        std::vector<V> s(C);
        for (int64_t c = 0; c < C; ++c) { s[c].x = 0; s[c].y = 0; s[c].z = 0; s[c].w = host_bed[c]; }
        d_state.upload(s, stream);
        std::vector<real> z(C, real(0));
        d_wet.upload(z, stream);
        if (P.compensated) {
            std::vector<V> zc(C);
            for (auto& v : zc) { v.x = 0; v.y = 0; v.z = 0; v.w = 0; }
            d_comp.upload(zc, stream);
        }
        // default forcing: no rain multiplier effect (1), no infiltration
        std::vector<real> ones(C, real(1));
        d_rainm.upload(ones, stream);
        d_f0.upload(z, stream); d_fc.upload(z, stream); d_k.upload(z, stream);
        // launch geometry and reduction scratch
        g_edge = (int)std::max<int64_t>(1, (E + bs - 1) / bs);
        g_cell = (int)std::max<int64_t>(1, (C + bs - 1) / bs);
        g_begin = (int)std::min<int64_t>(512, std::max<int64_t>(1, (C + 1023) / 1024));
        d_part_begin.alloc(g_begin);
        d_part_flux.alloc(2 * (size_t)g_edge);
        d_part_upd.alloc(2 * (size_t)g_cell);
        d_part_fin.alloc(1024);
        if (N == 0) ensure_empty_drainage();
        FH_CHECK(hipStreamSynchronize(stream));
        surface_ready = true;
    }

    void ensure_empty_drainage() {}

    void upload_coupling(int64_t N_, int64_t I_, int64_t S_, const int32_t* icell, const int32_t* inode,
                         const int32_t* nsc, const float* inv, const float* As, const float* dmax,
                         const float* capq, const float* surq, const int32_t* scell, const int32_t* snode,
                         const float* sw, int32_t snorm) override {
        need_surface();
        drop_graph();
        drain.upload_coupling(*this, N_, I_, S_, icell, inode, nsc, inv, As, dmax, capq, surq, scell, snode, sw, snorm);
        N = N_;
        I = I_;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void upload_pipe(int64_t L_, const int32_t* from, const int32_t* to, const float* len, const float* n,
                     const float* cap, const float* inv, const float* D, const float* minor, const int32_t* geom,
                     const float* rw, const float* rh, const float* oa, const float* oc, const float* otw,
                     const float* ocap) override {
        need_surface();
        drop_graph();
        drain.upload_pipe(*this, L_, from, to, len, n, cap, inv, D, minor, geom, rw, rh, oa, oc, otw, ocap);
        L = L_;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void upload_recharge(int64_t R_, const int32_t* cell, const float* S, const float* cc, const float* rc,
                         const float* stop) override {
        need_surface();
        drop_graph();
        drain.upload_recharge(*this, R_, cell, S, cc, rc, stop);
        R = R_;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void set_forcing(int64_t nr, const float* rs, const float* re, const float* rr, const float* m,
                     const float* f0, const float* fc, const float* k, int64_t nt, const float* et,
                     int64_t nrcv, const int32_t* rcv, const float* q) override {
        need_surface();
        drop_graph();
        if (nr < 0 || nt < 0 || nrcv < 0) throw ValueError{"negative forcing sizes"};
        if (nr && (!rs || !re || !rr)) throw ValueError{"rain interval arrays must not be NULL"};
        d_rain_s.upload(rs, (size_t)nr, stream);
        d_rain_e.upload(re, (size_t)nr, stream);
        d_rain_r.upload(rr, (size_t)nr, stream);
        n_rain = nr;
        std::vector<real> hm(C), hf0(C), hfc(C), hk(C);
        for (int64_t c = 0; c < C; ++c) {
            // multiplier folded with the active mask (rain_depth = rate * m * dt * active)
            hm[c] = host_active[c] ? (real)(m ? m[c] : 1.0f) : real(0);
            hf0[c] = (real)(f0 ? f0[c] : 0.0f);
            hfc[c] = (real)(fc ? fc[c] : 0.0f);
            hk[c] = (real)(k ? k[c] : 0.0f);
        }
        d_rainm.upload(hm, stream);
        d_f0.upload(hf0, stream); d_fc.upload(hfc, stream); d_k.upload(hk, stream);
        drain.set_external(*this, nt, et, nrcv, rcv, q);
        n_ext_times = nrcv > 0 ? nt : 0;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void set_state(const float* h, const float* hu, const float* hv, const float* wet, const float* vol,
                   const float* q, const float* rs) override {
        need_surface();
        std::vector<V> s(C);
        d_state.download(s.data(), C, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        for (int64_t c = 0; c < C; ++c) {
            if (h) s[c].x = (real)h[c];
            if (hu) s[c].y = (real)hu[c];
            if (hv) s[c].z = (real)hv[c];
        }
        d_state.upload(s, stream);
        if (wet) {
            std::vector<real> w(wet, wet + C);
            d_wet.upload(w, stream);
        }
        if (P.compensated && (h || hu || hv)) {
            std::vector<V> zc(C);
            for (auto& v : zc) { v.x = 0; v.y = 0; v.z = 0; v.w = 0; }
            d_comp.upload(zc, stream);
        }
        drain.set_state(*this, vol, q, rs);
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void finalize_pending() {
        fh::k_finalize_wet<real><<<g_fin(), 256, 0, stream>>>(ctl, kp, (int)C, d_state.p, d_wet.p);
        fh::k_clear_pending<<<1, 1, 0, stream>>>(ctl);
    }
    int g_fin() const { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, (C + 255) / 256)); }

    void get_state(float* h, float* hu, float* hv, float* wet, float* vol, float* q, float* rs) override {
        need_surface();
        std::vector<V> s(C);
        d_state.download(s.data(), C, stream);
        std::vector<real> w(C);
        d_wet.download(w.data(), C, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        for (int64_t c = 0; c < C; ++c) {
            if (h) h[c] = (float)s[c].x;
            if (hu) hu[c] = (float)s[c].y;
            if (hv) hv[c] = (float)s[c].z;
            if (wet) wet[c] = (float)w[c];
        }
        drain.get_state(*this, vol, q, rs);
    }

    void get_state_f64(double* h, double* hu, double* hv) override {
        need_surface();
        std::vector<V> s(C);
        d_state.download(s.data(), C, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        for (int64_t c = 0; c < C; ++c) {
            if (h) h[c] = (double)s[c].x;
            if (hu) hu[c] = (double)s[c].y;
            if (hv) hv[c] = (double)s[c].z;
        }
    }

    void pull_ctl() {
        FH_CHECK(hipMemcpyAsync(ctl_h, ctl, sizeof(fh::Ctl), hipMemcpyDeviceToHost, stream));
        FH_CHECK(hipStreamSynchronize(stream));
    }
    void push_ctl() {
        FH_CHECK(hipMemcpyAsync(ctl, ctl_h, sizeof(fh::Ctl), hipMemcpyHostToDevice, stream));
        FH_CHECK(hipStreamSynchronize(stream));
    }

    void set_clock(double t, double cached, int64_t step) override {
        pull_ctl();
        ctl_h->t = t;
        ctl_h->cached_cfl = (double)(real)cached;
        ctl_h->step = step;
        ctl_h->eval_next = (P.fixed_dt <= 0.0 && (step % kp.interval) == 0) ? 1 : 0;
        push_ctl();
    }

    void get_clock(fh_clock* c) override {
        pull_ctl();
        c->time_s = ctl_h->t;
        c->cached_dt_s = ctl_h->cached_cfl;
        c->last_dt_s = ctl_h->dt;
        c->global_step = ctl_h->step;
        c->interval_steps = ctl_h->interval_step;
        c->cfl_evaluations = ctl_h->cfl_evals;
    }

    void set_ledger(const double* l) override {
        pull_ctl();
        for (int k = 0; k < fh::kLedgerTerms; ++k) ctl_h->ledger[k] = l ? l[k] : 0.0;
        push_ctl();
    }
    void get_ledger(double* l) override {
        pull_ctl();
        for (int k = 0; k < fh::kLedgerTerms; ++k) l[k] = ctl_h->ledger[k];
    }

    // ------------------------------------------------------------------ stepping
    void launch_step() {
        fh::k_begin<real><<<g_begin, 256, 0, stream>>>(ctl, kp, d_state.p, d_lc.p, (int)C, d_part_begin.p,
                                                      d_rain_s.p, d_rain_e.p, d_rain_r.p, (int)n_rain,
                                                      drain.ext_times_ptr(), (int)n_ext_times);
        if (E > 0)
            fh::k_flux<real><<<g_edge, bs, 0, stream>>>(ctl, kp, (int)E_int, (int)E, d_el.p, d_er.p, d_nx.p,
                                                       d_ny.p, d_len.p, d_state.p, d_rec.p, d_recR.p,
                                                       d_part_flux.p);
        if (P.compensated)
            fh::k_update<real, true><<<g_cell, bs, 0, stream>>>(ctl, kp, (int)C, d_ce0.p, d_ce1.p, d_ce2.p,
                                                               d_rec.p, d_recR.p, d_state.p, d_comp.p, d_wet.p,
                                                               d_area.p, d_n2.p, d_rainm.p, d_f0.p, d_fc.p,
                                                               d_k.p, d_part_upd.p);
        else
            fh::k_update<real, false><<<g_cell, bs, 0, stream>>>(ctl, kp, (int)C, d_ce0.p, d_ce1.p, d_ce2.p,
                                                                d_rec.p, d_recR.p, d_state.p, d_comp.p, d_wet.p,
                                                                d_area.p, d_n2.p, d_rainm.p, d_f0.p, d_fc.p,
                                                                d_k.p, d_part_upd.p);
        drain.launch(*this);
    }

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }

    void build_graph(int k) {
        drop_graph();
        FH_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < k; ++i) launch_step();
        FH_CHECK(hipStreamEndCapture(stream, &graph));
        FH_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_steps = k;
    }

    void launch_chunk(int k) {
        if (P.use_graph) {
            if (!gexec || graph_steps != k) build_graph(k);
            FH_CHECK(hipGraphLaunch(gexec, stream));
        } else {
            for (int i = 0; i < k; ++i) launch_step();
        }
    }

    bool check_finite() {
        fh::k_finite<real><<<g_fin(), 256, 0, stream>>>(ctl, (int)C, d_state.p, (int)N, drain.volume_ptr(),
                                                        d_part_fin.p);
        pull_ctl();
        return ctl_h->nonfinite == 0;
    }

    int advance_to(double target, double final_t, int64_t max_steps) override {
        need_surface();
        if (max_steps <= 0) throw ValueError{"max_steps must be positive"};
        fh::k_set_target<<<1, 1, 0, stream>>>(ctl, target, final_t, max_steps);
        const int K = std::max(1, P.steps_per_launch);
        pull_ctl();
        int64_t done = 0;
        while (ctl_h->t < target && ctl_h->interval_step < max_steps) {
            // launch several chunks when far from the target (extra steps are no-ops)
            const double dt_est = std::max(ctl_h->dt > 0 ? ctl_h->dt : 1e-3, 1e-6);
            const double remaining = (target - ctl_h->t) / dt_est;
            int chunks = (int)std::min<double>(8.0, std::max(1.0, std::floor(remaining / K)));
            const int64_t budget_left = max_steps - ctl_h->interval_step;
            chunks = (int)std::max<int64_t>(1, std::min<int64_t>(chunks, (budget_left + K - 1) / K));
            for (int i = 0; i < chunks; ++i) launch_chunk(K);
            pull_ctl();
            done += (int64_t)chunks * K;
        }
        finalize_pending();
        const bool finite = check_finite();
        if (!finite) return FH_ERR_SIMULATION;
        if (ctl_h->t < target) return FH_ERR_BUDGET;
        return FH_OK;
    }

    void step_once(double dt, double rain, const float* q, double* diag) override {
        need_surface();
        if (!(dt > 0.0)) throw ValueError{"timestep_s must be positive."};
        drain.set_external_override(*this, q);
        fh::k_set_target<<<1, 1, 0, stream>>>(ctl, INFINITY, INFINITY, 1);
        fh::k_set_override<<<1, 1, 0, stream>>>(ctl, 1, dt, rain);
        launch_step();
        fh::k_set_override<<<1, 1, 0, stream>>>(ctl, 0, 0.0, 0.0);
        finalize_pending();
        pull_ctl();
        drain.clear_external_override(*this);
        // restore CFL cadence for the next policy step
        ctl_h->eval_next = (P.fixed_dt <= 0.0 && (ctl_h->step % kp.interval) == 0) ? 1 : 0;
        push_ctl();
        if (diag)
            for (int k = 0; k < fh::kLedgerTerms; ++k) diag[k] = ctl_h->diag[k];
    }

    double stable_timestep() override {
        need_surface();
        // run the device reduction of k_begin on an inactive step (target = -inf):
        // it records the CFL minimum and leaves the clock untouched
        pull_ctl();
        fh::Ctl saved = *ctl_h;
        ctl_h->eval_next = 1;
        ctl_h->override_dt = 0;
        ctl_h->target = -INFINITY;
        push_ctl();
        fh::k_begin<real><<<g_begin, 256, 0, stream>>>(ctl, kp, d_state.p, d_lc.p, (int)C, d_part_begin.p,
                                                      d_rain_s.p, d_rain_e.p, d_rain_r.p, 0, nullptr, 0);
        pull_ctl();
        const double r = ctl_h->cfl_min;
        *ctl_h = saved;
        ctl_h->cfl_min = r;
        push_ctl();
        return r;
    }

    std::vector<KernelTiming> profile(int steps) override;

    int64_t device_bytes() const override {
        int64_t b = 0;
        b += d_el.bytes() + d_er.bytes() + d_nx.bytes() + d_ny.bytes() + d_len.bytes();
        b += d_rec.bytes() + d_recR.bytes() + d_state.bytes() + d_wet.bytes() + d_comp.bytes();
        b += d_area.bytes() + d_n2.bytes() + d_lc.bytes() + d_rainm.bytes() + d_f0.bytes() + d_fc.bytes() + d_k.bytes();
        b += d_ce0.bytes() + d_ce1.bytes() + d_ce2.bytes();
        b += drain.device_bytes();
        return b;
    }

    void need_surface() const {
        if (!surface_ready) throw ValueError{"fh_upload_surface must be called first"};
    }

    // ------------------------------------------------------------------ members
    fh_params P;
    fh::KP<real> kp{};
    hipStream_t stream = nullptr;
    fh::Ctl* ctl = nullptr;
    fh::Ctl* ctl_h = nullptr;
    int bs = 256;
    int g_edge = 1, g_cell = 1, g_begin = 1;
    int64_t C = 0, E = 0, E_int = 0, N = 0, I = 0, L = 0, R = 0, n_rain = 0, n_ext_times = 0;
    bool surface_ready = false;
    std::vector<real> host_bed;
    std::vector<uint8_t> host_active;
    DevArray<int32_t> d_el, d_er, d_ce0, d_ce1, d_ce2;
    DevArray<real> d_nx, d_ny, d_len, d_recR, d_wet, d_area, d_n2, d_lc, d_rainm, d_f0, d_fc, d_k;
    DevArray<V> d_rec, d_state, d_comp;
    DevArray<double> d_part_begin, d_part_flux, d_part_upd, d_part_fin;
    DevArray<float> d_rain_s, d_rain_e, d_rain_r;
    fh::Drainage<real> drain;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
};

template <typename real>
std::vector<KernelTiming> Engine<real>::profile(int steps) {
    need_surface();
    // time each kernel of the step sequence with events on the context stream
    std::vector<KernelTiming> out;
    auto timed = [&](const char* name, double bytes, auto&& fn) {
        hipEvent_t a, b;
        FH_CHECK(hipEventCreate(&a));
        FH_CHECK(hipEventCreate(&b));
        FH_CHECK(hipEventRecord(a, stream));
        fn();
        FH_CHECK(hipEventRecord(b, stream));
        FH_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        FH_CHECK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        auto it = std::find_if(out.begin(), out.end(), [&](const KernelTiming& k) { return k.name == name; });
        if (it == out.end()) {
            out.push_back(KernelTiming{name, 0.0, bytes, 0});
            it = out.end() - 1;
        }
        it->total_ms += ms;
        it->launches += 1;
    };
    const double R4 = sizeof(real), I4 = 4.0;
    // algorithmic bytes (each required element once per step)
    const double flux_bytes = (double)E * (2 * I4 + 3 * R4) + (double)C * 5 * R4 + (double)E * 5 * R4;
    const double upd_bytes = (double)E * 5 * R4 + (double)C * (3 * I4 + 6 * R4 + 3 * R4 + 4 * R4 + 2 * R4) +
                             (P.compensated ? (double)C * 6 * R4 : 0.0);
    fh::k_set_target<<<1, 1, 0, stream>>>(ctl, INFINITY, INFINITY, (int64_t)steps + 8);
    for (int i = 0; i < steps; ++i) {
        timed("begin", (double)C * 4 * R4 / kp.interval, [&] {
            fh::k_begin<real><<<g_begin, 256, 0, stream>>>(ctl, kp, d_state.p, d_lc.p, (int)C, d_part_begin.p,
                                                          d_rain_s.p, d_rain_e.p, d_rain_r.p, (int)n_rain,
                                                          drain.ext_times_ptr(), (int)n_ext_times);
        });
        if (E > 0)
            timed("flux", flux_bytes, [&] {
                fh::k_flux<real><<<g_edge, bs, 0, stream>>>(ctl, kp, (int)E_int, (int)E, d_el.p, d_er.p, d_nx.p,
                                                           d_ny.p, d_len.p, d_state.p, d_rec.p, d_recR.p,
                                                           d_part_flux.p);
            });
        timed("update", upd_bytes, [&] {
            if (P.compensated)
                fh::k_update<real, true><<<g_cell, bs, 0, stream>>>(ctl, kp, (int)C, d_ce0.p, d_ce1.p, d_ce2.p,
                                                                   d_rec.p, d_recR.p, d_state.p, d_comp.p,
                                                                   d_wet.p, d_area.p, d_n2.p, d_rainm.p, d_f0.p,
                                                                   d_fc.p, d_k.p, d_part_upd.p);
            else
                fh::k_update<real, false><<<g_cell, bs, 0, stream>>>(ctl, kp, (int)C, d_ce0.p, d_ce1.p, d_ce2.p,
                                                                    d_rec.p, d_recR.p, d_state.p, d_comp.p,
                                                                    d_wet.p, d_area.p, d_n2.p, d_rainm.p, d_f0.p,
                                                                    d_fc.p, d_k.p, d_part_upd.p);
        });
        drain.profile(*this, timed);
    }
    finalize_pending();
    FH_CHECK(hipStreamSynchronize(stream));
    return out;
}

struct Ctx {
    std::unique_ptr<EngineBase> eng;
    std::string err;
};

}  // namespace

struct fh_ctx {
    Ctx c;
};

#define FH_GUARD(ctx, body)                                                        \
    do {                                                                           \
        if (!(ctx)) { g_error = "NULL context"; return FH_ERR_VALUE; }             \
        try {                                                                      \
            body;                                                                  \
        } catch (const ValueError& v) {                                            \
            (ctx)->c.err = v.msg;                                                  \
            return FH_ERR_VALUE;                                                   \
        } catch (const HipError& h) {                                              \
            (ctx)->c.err = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.where; \
            return FH_ERR_RUNTIME;                                                 \
        } catch (const std::exception& x) {                                        \
            (ctx)->c.err = x.what();                                               \
            return FH_ERR_RUNTIME;                                                 \
        }                                                                          \
    } while (0)

extern "C" {

void fh_default_params(fh_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->gravity = 9.80665;
    p->min_depth = 1.0e-5;
    p->cfl = 0.85;
    p->reuse_factor = 0.8;
    p->max_dt = 1.0;
    p->dt_floor = 1.0e-3;
    p->max_speed = 5.0;
    p->inlet_orifice_q = 0.62 * 0.04;
    p->slot_wave_speed = 30.0;
    p->fixed_dt = 0.0;
    p->cfl_interval = 32;
    p->compensated = 0;
    p->precision = 32;
    p->use_graph = 1;
    p->steps_per_launch = 32;
    p->block_size = 256;
    p->xcd_remap = 1;
    p->device = 0;
}

const char* fh_version(void) { return "floodhip 0.1.0 (gfx950)"; }

int fh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* fh_global_error(void) { return g_error.c_str(); }

fh_ctx* fh_create(const fh_params* params) {
    fh_params p;
    if (params) p = *params; else fh_default_params(&p);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_error = "no HIP device is visible (BackendUnavailableError)";
        return nullptr;
    }
    if (p.device < 0 || p.device >= n) {
        g_error = "device ordinal out of range";
        return nullptr;
    }
    if (p.precision != 32 && p.precision != 64) {
        g_error = "precision must be 32 or 64";
        return nullptr;
    }
    try {
        auto* ctx = new fh_ctx();
        if (p.precision == 64) ctx->c.eng.reset(new Engine<double>(p));
        else ctx->c.eng.reset(new Engine<float>(p));
        return ctx;
    } catch (const HipError& h) {
        g_error = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.where;
    } catch (const std::exception& x) {
        g_error = x.what();
    }
    return nullptr;
}

void fh_destroy(fh_ctx* ctx) { delete ctx; }

const char* fh_last_error(const fh_ctx* ctx) { return ctx ? ctx->c.err.c_str() : g_error.c_str(); }

int fh_upload_surface(fh_ctx* ctx, int64_t n_cells, int64_t n_edges, const float* bed, const float* manning_n,
                      const float* cell_area, const float* lc, const int32_t* edge_left,
                      const int32_t* edge_right, const float* edge_nx, const float* edge_ny,
                      const float* edge_length, const int32_t* boundary_code, const uint8_t* active) {
    FH_GUARD(ctx, ctx->c.eng->upload_surface(n_cells, n_edges, bed, manning_n, cell_area, lc, edge_left,
                                             edge_right, edge_nx, edge_ny, edge_length, boundary_code, active));
    return FH_OK;
}

int fh_upload_coupling(fh_ctx* ctx, int64_t n_nodes, int64_t n_inlets, int64_t n_entries,
                       const int32_t* inlet_cell, const int32_t* inlet_node, const int32_t* node_surface_cell,
                       const float* node_invert, const float* node_storage_area, const float* node_max_depth,
                       const float* inlet_capture_capacity, const float* inlet_surcharge_capacity,
                       const int32_t* surcharge_cell, const int32_t* surcharge_node, const float* surcharge_weight,
                       int32_t surcharge_weight_normalized) {
    FH_GUARD(ctx, ctx->c.eng->upload_coupling(n_nodes, n_inlets, n_entries, inlet_cell, inlet_node,
                                              node_surface_cell, node_invert, node_storage_area, node_max_depth,
                                              inlet_capture_capacity, inlet_surcharge_capacity, surcharge_cell,
                                              surcharge_node, surcharge_weight, surcharge_weight_normalized));
    return FH_OK;
}

int fh_upload_pipe(fh_ctx* ctx, int64_t n_links, const int32_t* link_from, const int32_t* link_to,
                   const float* link_length, const float* link_manning, const float* link_capacity,
                   const float* link_invert, const float* link_diameter, const float* link_minor_loss,
                   const int32_t* link_geometry_code, const float* link_rect_width, const float* link_rect_height,
                   const float* outfall_area, const float* outfall_coefficient, const float* outfall_tailwater,
                   const float* outfall_capacity) {
    FH_GUARD(ctx, ctx->c.eng->upload_pipe(n_links, link_from, link_to, link_length, link_manning, link_capacity,
                                          link_invert, link_diameter, link_minor_loss, link_geometry_code,
                                          link_rect_width, link_rect_height, outfall_area, outfall_coefficient,
                                          outfall_tailwater, outfall_capacity));
    return FH_OK;
}

int fh_upload_recharge(fh_ctx* ctx, int64_t n_rows, const int32_t* source_cell, const float* storage_capacity,
                       const float* capture_capacity, const float* recharge_capacity, const float* stop_depth) {
    FH_GUARD(ctx, ctx->c.eng->upload_recharge(n_rows, source_cell, storage_capacity, capture_capacity,
                                              recharge_capacity, stop_depth));
    return FH_OK;
}

int fh_set_forcing(fh_ctx* ctx, int64_t n_intervals, const float* rain_start_s, const float* rain_end_s,
                   const float* rain_rate_mps, const float* rainfall_multiplier, const float* horton_f0_mps,
                   const float* horton_fc_mps, const float* horton_k_per_s, int64_t n_times,
                   const float* external_times_s, int64_t n_receivers, const int32_t* external_receiver_cell,
                   const float* external_q_m3ps) {
    FH_GUARD(ctx, ctx->c.eng->set_forcing(n_intervals, rain_start_s, rain_end_s, rain_rate_mps, rainfall_multiplier,
                                          horton_f0_mps, horton_fc_mps, horton_k_per_s, n_times, external_times_s,
                                          n_receivers, external_receiver_cell, external_q_m3ps));
    return FH_OK;
}

int fh_set_state(fh_ctx* ctx, const float* h, const float* hu, const float* hv, const float* wet_time_s,
                 const float* node_volume, const float* link_flow, const float* recharge_storage) {
    FH_GUARD(ctx, ctx->c.eng->set_state(h, hu, hv, wet_time_s, node_volume, link_flow, recharge_storage));
    return FH_OK;
}

int fh_get_state(fh_ctx* ctx, float* h, float* hu, float* hv, float* wet_time_s, float* node_volume,
                 float* link_flow, float* recharge_storage) {
    FH_GUARD(ctx, ctx->c.eng->get_state(h, hu, hv, wet_time_s, node_volume, link_flow, recharge_storage));
    return FH_OK;
}

int fh_get_state_f64(fh_ctx* ctx, double* h, double* hu, double* hv) {
    FH_GUARD(ctx, ctx->c.eng->get_state_f64(h, hu, hv));
    return FH_OK;
}

int fh_set_clock(fh_ctx* ctx, double time_s, double cached_dt_s, int64_t global_step) {
    FH_GUARD(ctx, ctx->c.eng->set_clock(time_s, cached_dt_s, global_step));
    return FH_OK;
}

int fh_get_clock(fh_ctx* ctx, fh_clock* out) {
    if (!out) return FH_ERR_VALUE;
    FH_GUARD(ctx, ctx->c.eng->get_clock(out));
    return FH_OK;
}

int fh_set_ledger(fh_ctx* ctx, const double* ledger13) {
    FH_GUARD(ctx, ctx->c.eng->set_ledger(ledger13));
    return FH_OK;
}

int fh_get_ledger(fh_ctx* ctx, double* ledger13) {
    if (!ledger13) return FH_ERR_VALUE;
    FH_GUARD(ctx, ctx->c.eng->get_ledger(ledger13));
    return FH_OK;
}

int fh_advance_to(fh_ctx* ctx, double target_s, double final_s, int64_t max_steps, fh_clock* clock_out,
                  double* ledger_out) {
    int status = FH_OK;
    FH_GUARD(ctx, {
        status = ctx->c.eng->advance_to(target_s, final_s, max_steps);
        if (clock_out) ctx->c.eng->get_clock(clock_out);
        if (ledger_out) ctx->c.eng->get_ledger(ledger_out);
        if (status == FH_ERR_SIMULATION) ctx->c.err = "non-finite surface depth or drainage storage";
        if (status == FH_ERR_BUDGET) ctx->c.err = "step budget exhausted before the target time";
    });
    return status;
}

int fh_step(fh_ctx* ctx, double dt, double rain_rate_mps, const float* external_q_m3ps, double* diag_out) {
    FH_GUARD(ctx, ctx->c.eng->step_once(dt, rain_rate_mps, external_q_m3ps, diag_out));
    return FH_OK;
}

int fh_stable_timestep(fh_ctx* ctx, double* dt_out) {
    if (!dt_out) return FH_ERR_VALUE;
    FH_GUARD(ctx, *dt_out = ctx->c.eng->stable_timestep());
    return FH_OK;
}

int fh_profile_kernels(fh_ctx* ctx, int32_t steps, int32_t cap, int32_t* count, char* names, double* avg_ms,
                       double* bytes, int64_t* launches) {
    FH_GUARD(ctx, {
        auto t = ctx->c.eng->profile(steps);
        int n = std::min<int>(cap, (int)t.size());
        for (int i = 0; i < n; ++i) {
            std::memset(names + 32 * i, 0, 32);
            std::strncpy(names + 32 * i, t[i].name.c_str(), 31);
            avg_ms[i] = t[i].launches ? t[i].total_ms / (double)t[i].launches : 0.0;
            bytes[i] = t[i].bytes;
            launches[i] = t[i].launches;
        }
        *count = n;
    });
    return FH_OK;
}

int64_t fh_device_bytes(const fh_ctx* ctx) { return ctx ? ctx->c.eng->device_bytes() : 0; }

}  // extern "C"
