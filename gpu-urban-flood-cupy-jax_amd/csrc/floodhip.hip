// floodhip: host engine and C ABI (see include/floodhip.h).
//
// An Engine<real> holds a batch of M members (M = 1 for the plain solver):
// independent scenarios, each with its own mesh, drainage network, forcing,
// clock, dt and FP64 ledger.  Uploads stage each member on the host; pack()
// concatenates the members into device arrays member by member, each member's
// range of every index space starting on a tile boundary, and builds the
// block -> member tables the kernels use.  A step is one fixed sequence of
// kernel launches covering every member; fh_advance_to replays captured
// hipGraphs of steps_per_launch steps and reads the control blocks back once
// per replay batch.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <memory>
#include <numeric>
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
    // uploads v (at least one element is always allocated so kernels get a valid pointer)
    // An empty table still gets one zeroed element, so kernels may load entry 0
    // of any table ahead of their row tests (and use it as an index).
    void set(const std::vector<T>& v, hipStream_t s) {
        const size_t cnt = std::max<size_t>(v.size(), 1);
        if (cnt != n) alloc(cnt);
        if (!v.empty()) FH_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        else FH_CHECK(hipMemsetAsync(p, 0, sizeof(T), s));
    }
    void put(const T* host, size_t offset, size_t count, hipStream_t s) {
        if (count) FH_CHECK(hipMemcpyAsync(p + offset, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void get(T* host, size_t offset, size_t count, hipStream_t s) const {
        if (count) FH_CHECK(hipMemcpyAsync(host, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    size_t bytes() const { return n * sizeof(T); }
    ~DevArray() { release(); }
};

// One index space of the batch: members concatenated on tile boundaries.
struct Layout {
    int tile = 256;
    std::vector<int> base, count;              // per member (in pack order)
    std::vector<int> blk_member, mem_blk0, mem_nblk, mem_end;
    int total = 0, blocks = 0;
    void build(const std::vector<int64_t>& counts, int tile_) {
        tile = tile_;
        const int M = (int)counts.size();
        base.assign(M, 0); count.assign(M, 0);
        mem_blk0.assign(M, 0); mem_nblk.assign(M, 0); mem_end.assign(M, 0);
        blk_member.clear();
        int64_t pos = 0;
        for (int m = 0; m < M; ++m) {
            const int64_t nb = (counts[m] + tile - 1) / tile;
            base[m] = (int)pos;
            count[m] = (int)counts[m];
            mem_blk0[m] = (int)(pos / tile);
            mem_nblk[m] = (int)nb;
            mem_end[m] = (int)(pos + counts[m]);
            for (int64_t b = 0; b < nb; ++b) blk_member.push_back(m);
            pos += nb * tile;
        }
        if (pos > INT32_MAX / 2) throw ValueError{"batch too large for 32-bit indices"};
        total = (int)pos;
        blocks = (int)(pos / tile);
    }
};

// Permute a per-member table from pack position order to member id order.
template <typename T>
std::vector<T> by_id(const std::vector<T>& by_pos, const std::vector<int>& order) {
    std::vector<T> out(by_pos.size());
    for (size_t i = 0; i < by_pos.size(); ++i) out[(size_t)order[i]] = by_pos[i];
    return out;
}

// Device copy of a Layout; kernels index every per-member table by member id.
struct DevSeg {
    DevArray<int> blk_member, mem_blk0, mem_nblk, mem_end;
    void set(const Layout& L, const std::vector<int>& order, hipStream_t s) {
        std::vector<int> bm(L.blk_member.size());
        for (size_t b = 0; b < bm.size(); ++b) bm[b] = order[(size_t)L.blk_member[b]];
        if (bm.empty()) bm.push_back(0);
        blk_member.set(bm, s);
        mem_blk0.set(by_id(L.mem_blk0, order), s);
        mem_nblk.set(by_id(L.mem_nblk, order), s);
        mem_end.set(by_id(L.mem_end, order), s);
    }
    fh::Seg seg() const { return fh::Seg{blk_member.p, mem_blk0.p, mem_nblk.p, mem_end.p}; }
};

struct KernelTiming {
    std::string name;
    double total_ms = 0.0;
    double bytes = 0.0;
    int64_t launches = 0;
};

// Counting-sort CSR: rows[i] in [0, nrows) for items i (negative rows skipped);
// items of a row stay in ascending item order.
void build_csr(const std::vector<int>& rows, int nrows, std::vector<int>& ptr, std::vector<int>& idx) {
    ptr.assign((size_t)nrows + 1, 0);
    for (int r : rows)
        if (r >= 0) ptr[(size_t)r + 1]++;
    for (int r = 0; r < nrows; ++r) ptr[(size_t)r + 1] += ptr[r];
    idx.assign((size_t)ptr[nrows], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < rows.size(); ++i)
        if (rows[i] >= 0) idx[(size_t)fill[rows[i]]++] = (int)i;
}

template <typename real>
std::vector<real> conv(const float* src, int64_t n, float dflt = 0.0f) {
    std::vector<real> v((size_t)std::max<int64_t>(n, 0));
    for (int64_t i = 0; i < n; ++i) v[(size_t)i] = (real)(src ? src[i] : dflt);
    return v;
}
std::vector<int> ints(const int32_t* src, int64_t n, int dflt = -1) {
    std::vector<int> v((size_t)std::max<int64_t>(n, 0));
    for (int64_t i = 0; i < n; ++i) v[(size_t)i] = src ? src[i] : dflt;
    return v;
}

// Host staging of one member (local indices).
template <typename real>
struct Member {
    bool has_surface = false;
    int64_t C = 0, E = 0, E_int = 0;
    std::vector<int> el, er;
    std::vector<real> nx, ny, len;
    std::vector<int> ce0, ce1, ce2;     // local edge positions, encoded
    std::vector<real> area, n2, n1, lc, bed;
    std::vector<uint8_t> active;
    std::vector<real> rainm, f0, fc, hk;
    std::vector<float> rain_s, rain_e, rain_r;
    // stage boundary: the staged positions of the code-3 edges and their caller
    // edge ids; stage = one value per caller edge (empty: no stage values)
    int64_t E_caller = 0;
    std::vector<int> st_pos, st_edge;
    std::vector<char> st_dead;   // the code-3 edge's cell is inactive
    std::vector<float> stage;
    // drainage
    int64_t N = 0, I = 0, L = 0, S = 0, R = 0, Rcv = 0, T = 0;
    std::vector<real> inv, As, dmax, oarea, ocoef, otw, ocap;
    std::vector<int> nsc;
    std::vector<int> icell, inode;
    std::vector<real> capq, surq;
    std::vector<int> lfrom, lto, geom;
    std::vector<real> llen, man, lcap, linv, D, minor, rw, rh;
    std::vector<int> scell, snode;
    std::vector<real> sw;
    int snorm = 0;
    std::vector<int> rcell;
    std::vector<real> rS, rcc, rrc, rstop;
    std::vector<float> ext_t, ext_q;
    std::vector<int> rcv;
    // state (host copy; authoritative only while the batch is not packed)
    std::vector<real> h, hu, hv, wet, V, Q, rstore;
    fh::Ctl ctl{};
    // device layout (member-local; empty = as given): cpos[c] = position of the
    // caller's cell c, epos[i] = position of staged edge i, npos / lpos = positions
    // of the caller's nodes / links.  bfs caches the breadth-first cell order.
    std::vector<int> bfs, cpos, epos, npos, lpos;
    int cp(int c) const { return cpos.empty() ? c : cpos[(size_t)c]; }
    int np(int n) const { return npos.empty() ? n : npos[(size_t)n]; }
    int lp(int l) const { return lpos.empty() ? l : lpos[(size_t)l]; }
};

// Breadth-first (Cuthill-McKee) order of the cells over interior edges: each
// component starts at a pseudo-peripheral cell, neighbours are visited by
// increasing degree.  Cells that share an edge end up a front width apart, so
// the gathers of the flux and update stages stay within the L2 working set.
template <typename real>
void locality_order(Member<real>& m) {
    const int64_t C = m.C, Ei = m.E_int;
    std::vector<int> deg((size_t)C + 1, 0);
    for (int64_t e = 0; e < Ei; ++e) { deg[(size_t)m.el[e] + 1]++; deg[(size_t)m.er[e] + 1]++; }
    for (int64_t c = 0; c < C; ++c) deg[(size_t)c + 1] += deg[(size_t)c];
    std::vector<int> adj((size_t)deg[(size_t)C]);
    {
        std::vector<int> fill(deg.begin(), deg.end() - 1);
        for (int64_t e = 0; e < Ei; ++e) {
            adj[(size_t)fill[(size_t)m.el[e]]++] = m.er[e];
            adj[(size_t)fill[(size_t)m.er[e]]++] = m.el[e];
        }
    }
    auto degree = [&](int c) { return deg[(size_t)c + 1] - deg[(size_t)c]; };
    std::vector<int> order;
    order.reserve((size_t)C);
    std::vector<int> mark((size_t)C, 0);   // 0 unseen, else stamp of the BFS that saw it
    int stamp = 0;
    std::vector<int> queue;
    queue.reserve((size_t)C);
    std::vector<int> nb;
    // BFS from `root` over unplaced cells; returns the visit list (in queue)
    auto bfs = [&](int root, bool place) {
        ++stamp;
        queue.clear();
        queue.push_back(root);
        mark[(size_t)root] = place ? -1 : stamp;
        for (size_t q = 0; q < queue.size(); ++q) {
            const int c = queue[q];
            // neighbours by ascending degree, ties in adjacency order (a stable
            // insertion sort: a cell has a few neighbours, and std::stable_sort
            // would allocate a buffer per cell)
            nb.assign(adj.begin() + deg[(size_t)c], adj.begin() + deg[(size_t)c + 1]);
            for (size_t i = 1; i < nb.size(); ++i) {
                const int v = nb[i], dv = degree(v);
                size_t j = i;
                for (; j > 0 && degree(nb[j - 1]) > dv; --j) nb[j] = nb[j - 1];
                nb[j] = v;
            }
            for (int x : nb) {
                const int mk = mark[(size_t)x];
                if (mk == -1 || mk == (place ? -1 : stamp)) continue;
                mark[(size_t)x] = place ? -1 : stamp;
                queue.push_back(x);
            }
        }
    };
    for (int64_t c0 = 0; c0 < C; ++c0) {
        if (mark[(size_t)c0] == -1) continue;
        int root = (int)c0;
        // two sweeps towards a peripheral cell: the last cell reached, lowest degree among the last ones
        for (int sweep = 0; sweep < 2; ++sweep) {
            bfs(root, false);
            int best = queue.back();
            for (size_t q = queue.size(); q-- > 0 && queue.size() - q <= 64;)
                if (degree(queue[q]) < degree(best)) best = queue[q];
            root = best;
        }
        bfs(root, true);
        order.insert(order.end(), queue.begin(), queue.end());
    }
    m.bfs = std::move(order);
}

// Device positions: cells in breadth-first order; edges follow their lower
// cell position (interior block and ghost block each keep their place), ties
// in staged order.  Nodes follow the position of the first cell they exchange
// water with (their inlets' cells, else their surface cell; nodes with neither
// keep their order at the end) and links their lower end node, so the node- and
// link-side gathers of the drainage kernels walk the arrays in the same order
// as the cells they serve.
// Positions pos[i] (offset by base) of the items i = begin .. end-1 ordered by
// key(i) in [0, K), ties in index order: a counting sort, the order a stable
// comparison sort by key gives, in linear time.
template <class Key>
void positions_by_key(int64_t begin, int64_t end, int64_t K, int64_t base, Key key, std::vector<int>& pos) {
    std::vector<int64_t> start((size_t)K + 1, 0);
    for (int64_t i = begin; i < end; ++i) ++start[(size_t)key(i) + 1];
    for (int64_t k = 0; k < K; ++k) start[(size_t)k + 1] += start[(size_t)k];
    for (int64_t i = begin; i < end; ++i) pos[(size_t)i] = (int)(base + start[(size_t)key(i)]++);
}

template <typename real>
void assign_positions(Member<real>& m) {
    const int64_t C = m.C, Ei = m.E_int, E = m.E;
    m.cpos.assign((size_t)C, 0);
    for (int64_t i = 0; i < C; ++i) m.cpos[(size_t)m.bfs[(size_t)i]] = (int)i;
    // edges: by the lower new cell index (interior), by the new left cell (ghost)
    m.epos.assign((size_t)E, 0);
    positions_by_key(0, Ei, C, 0, [&](int64_t e) {
        return std::min(m.cpos[(size_t)m.el[e]], m.cpos[(size_t)m.er[e]]);
    }, m.epos);
    positions_by_key(Ei, E, C, Ei, [&](int64_t e) { return m.cpos[(size_t)m.el[e]]; }, m.epos);
    // nodes and links
    const int64_t N = m.N, L = m.L;
    std::vector<int64_t> nkey((size_t)N);
    for (int64_t n = 0; n < N; ++n) nkey[(size_t)n] = m.nsc[n] >= 0 ? m.cpos[(size_t)m.nsc[n]] : C + n;
    for (int64_t k = 0; k < m.I; ++k) {
        const int n = m.inode[k];
        nkey[(size_t)n] = std::min<int64_t>(nkey[(size_t)n], m.cpos[(size_t)m.icell[k]]);
    }
    m.npos.assign((size_t)N, 0);
    positions_by_key(0, N, C + N, 0, [&](int64_t n) { return nkey[(size_t)n]; }, m.npos);
    m.lpos.assign((size_t)L, 0);
    positions_by_key(0, L, std::max<int64_t>(N, 1), 0, [&](int64_t l) {
        return std::min(m.npos[(size_t)m.lfrom[l]], m.npos[(size_t)m.lto[l]]);
    }, m.lpos);
}

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual void select(int m) = 0;
    virtual int members() const = 0;
    virtual void upload_surface(int64_t C, int64_t E, const float* bed, const float* n, const float* area,
                                const float* lc, const int32_t* el, const int32_t* er, const float* nx,
                                const float* ny, const float* len, const int32_t* code, const uint8_t* active) = 0;
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
    virtual void set_boundary_stage(int64_t n, const float* stage) = 0;
    virtual void set_state(const float* h, const float* hu, const float* hv, const float* wet,
                           const float* vol, const float* q, const float* rs) = 0;
    virtual void get_state(float* h, float* hu, float* hv, float* wet, float* vol, float* q, float* rs) = 0;
    virtual void get_state_f64(double* h, double* hu, double* hv) = 0;
    virtual void get_drainage_f64(double* vol, double* q) = 0;
    virtual void set_clock(double t, double cached, int64_t step) = 0;
    virtual void get_clock(fh_clock* c) = 0;
    virtual void set_ledger(const double* l) = 0;
    virtual void get_ledger(double* l) = 0;
    virtual int advance_to(double target, double final_t, int64_t max_steps) = 0;
    virtual void step_once(double dt, double rain, const float* q, double* diag) = 0;
    virtual void dry_weather_step(double dt, const float* inflow, const float* tw, int cap, double* diag) = 0;
    virtual double stable_timestep() = 0;
    virtual std::vector<KernelTiming> profile(int steps) = 0;
    virtual int64_t device_bytes() const = 0;
};

template <typename real>
class Engine final : public EngineBase {
    typedef typename fh::V4<real>::type V;

public:
    Engine(const fh_params& prm, int members) : P(prm), M(members), mem((size_t)members) {
        FH_CHECK(hipSetDevice(P.device));
        FH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        FH_CHECK(hipMalloc((void**)&ctl, sizeof(fh::Ctl) * M));
        FH_CHECK(hipHostMalloc((void**)&ctl_h, sizeof(fh::Ctl) * M, hipHostMallocDefault));
        for (auto& m : mem) {
            std::memset(&m.ctl, 0, sizeof(fh::Ctl));
            const double inf = INFINITY;
            for (int q = 0; q < 2; ++q) {
                fh::Clock& ck = m.ctl.clk[q];
                ck.cached_cfl = 1.0;
                ck.eval = P.fixed_dt > 0.0 ? 0 : 1;
                ck.rain.next_end = -INFINITY;
                m.ctl.li_cfl_step[q] = -1;
            }
        }
        for (int m = 0; m < M; ++m) ctl_h[m] = mem[m].ctl;
        FH_CHECK(hipMemcpyAsync(ctl, ctl_h, sizeof(fh::Ctl) * M, hipMemcpyHostToDevice, stream));
        // the CFL minimum slots, all +inf
        {
            const double inf = INFINITY;
            unsigned long long ib;
            std::memcpy(&ib, &inf, sizeof(ib));
            cfl_inf.assign((size_t)M * 2 * fh::kCflSlots * fh::kCflStride, ib);
            d_cfl.set(cfl_inf, stream);
        }
        bs = std::max(64, std::min(256, P.block_size > 0 ? P.block_size : 256));
        bs = (bs / 64) * 64;
        ept = P.edges_per_thread == 2 ? 2 : (P.edges_per_thread == 4 ? 4 : 1);
        kp.g = (real)P.gravity;
        kp.eps = (real)P.min_depth;
        kp.half_g = (real)(0.5 * P.gravity);
        kp.two_g = (real)(2.0 * P.gravity);
        kp.cfl = (real)P.cfl;
        kp.max_speed = (real)P.max_speed;
        kp.tiny = (real)1.0e-12;
        kp.capture_q = (real)P.inlet_orifice_q;
        kp.pow43 = (real)(4.0 / 3.0);
        kp.pow43_dln2 = ((double)(float)(4.0 / 3.0) - 4.0 / 3.0) * M_LN2;
        kp.reuse = (real)P.reuse_factor;
        kp.slot_coef = (real)(P.slot_wave_speed * P.slot_wave_speed);
        kp.max_dt = P.max_dt;
        kp.dt_floor = P.dt_floor;
        kp.fixed_dt = P.fixed_dt > 0.0 ? (double)(real)P.fixed_dt : 0.0;
        kp.interval = std::max(1, P.cfl_interval);
        kp.g_d = P.gravity;
        if (P.surface_mode != 0 && P.surface_mode != 1) throw ValueError{"surface_mode must be 0 or 1"};
        li = P.surface_mode == 1;
        if (li) ept = 1;
        kp.remap = P.xcd_remap;
        kp.cfl_slots = d_cfl.p;
        FH_CHECK(hipStreamSynchronize(stream));
    }

    ~Engine() override {
        drop_graph();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ctl) (void)hipFree(ctl);
        if (ctl_h) (void)hipHostFree(ctl_h);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void select(int m) override {
        if (m < 0 || m >= M) throw ValueError{"member index out of range"};
        cur = m;
    }
    int members() const override { return M; }
    Member<real>& cm() { return mem[(size_t)cur]; }

    // ------------------------------------------------------------------ staging
    void touch() {
        if (packed && !dirty) pull_all();   // device state becomes the host copy before repacking
        dirty = true;
        drop_graph();
    }

    void upload_surface(int64_t C_, int64_t E_, const float* bed, const float* n, const float* area,
                        const float* lc_, const int32_t* el, const int32_t* er, const float* nx,
                        const float* ny, const float* len, const int32_t* code,
                        const uint8_t* active) override {
        if (C_ <= 0 || E_ < 0 || C_ > INT32_MAX / 4 || E_ > INT32_MAX / 4) throw ValueError{"surface sizes out of range"};
        if (!bed || !n || !area || !lc_ || (E_ && (!el || !er || !nx || !ny || !len || !code)))
            throw ValueError{"surface arrays must not be NULL"};
        for (int64_t e = 0; e < E_; ++e)
            if (el[e] < 0 || el[e] >= C_ || er[e] >= C_) throw ValueError{"edge cell index out of range"};
        touch();
        Member<real>& s = cm();
        auto act = [&](int64_t c) { return active ? active[c] != 0 : true; };
        std::vector<int64_t> interior, ghost;
        std::vector<int> ghost_kind;
        for (int64_t e = 0; e < E_; ++e) {
            const int32_t l = el[e], r = er[e];
            if (!act(l)) {
                // kept: a stage can flow into the inactive cell (kGhostDeadStage)
                if (r < 0 && code[e] == 3) { ghost.push_back(e); ghost_kind.push_back(fh::kGhostDeadOther); }
                continue;
            }
            const bool in = r >= 0;
            const bool ra = in && act(r);
            if (in && ra) { interior.push_back(e); continue; }
            const int32_t cd = (in && !ra) ? 0 : code[e];
            int kind;
            if (cd == 0 || cd == 2) kind = fh::kGhostReflect;
            else if (!in && cd == 4) kind = fh::kGhostOutflow;
            else if (cd == 1) kind = fh::kGhostCopy;
            else kind = fh::kGhostOther;   // code 3 turns into kGhostStage when stage values are set
            ghost.push_back(e);
            ghost_kind.push_back(kind);
        }
        s.C = C_;
        s.bfs.clear();
        s.cpos.clear();
        s.epos.clear();
        s.E_int = (int64_t)interior.size();
        s.E = s.E_int + (int64_t)ghost.size();
        std::vector<int32_t> pos_of((size_t)E_, -1);
        s.el.assign(s.E, 0); s.er.assign(s.E, 0);
        s.nx.assign(s.E, 0); s.ny.assign(s.E, 0); s.len.assign(s.E, 0);
        s.E_caller = E_;
        s.st_pos.clear(); s.st_edge.clear(); s.st_dead.clear(); s.stage.clear();
        for (int64_t i = 0; i < s.E; ++i) {
            const bool gh = i >= s.E_int;
            const int64_t e = gh ? ghost[i - s.E_int] : interior[i];
            pos_of[e] = (int32_t)i;
            s.el[i] = el[e];
            s.er[i] = gh ? ghost_kind[i - s.E_int] : er[e];
            s.nx[i] = (real)nx[e]; s.ny[i] = (real)ny[e]; s.len[i] = (real)len[e];
            if (gh && er[e] < 0 && code[e] == 3) {
                s.st_pos.push_back((int)i); s.st_edge.push_back((int)e); s.st_dead.push_back(act(el[e]) ? 0 : 1);
            }
        }
        // cell -> edge table: left-side entries (ascending original edge), then right-side
        std::vector<int> lent(C_ * 3, -1), rent(C_ * 3, -1);
        std::vector<uint8_t> nl(C_, 0), nr(C_, 0);
        for (int64_t e = 0; e < E_; ++e) {
            const int32_t p = pos_of[e];
            if (p < 0) continue;
            const int32_t l = el[e];
            if (nl[l] + nr[l] >= 3) throw ValueError{"a cell has more than three edges"};
            lent[l * 3 + nl[l]++] = p << 1;
            if (p < s.E_int) {
                const int32_t r = er[e];
                if (nl[r] + nr[r] >= 3) throw ValueError{"a cell has more than three edges"};
                rent[r * 3 + nr[r]++] = (p << 1) | 1;
            }
        }
        s.ce0.assign(C_, -1); s.ce1.assign(C_, -1); s.ce2.assign(C_, -1);
        for (int64_t c = 0; c < C_; ++c) {
            int t[3] = {-1, -1, -1};
            int k = 0;
            for (int i = 0; i < nl[c]; ++i) t[k++] = lent[c * 3 + i];
            for (int i = 0; i < nr[c]; ++i) t[k++] = rent[c * 3 + i];
            if (!act(c)) {
                // an inactive cell has only its kept code-3 boundary edges, in slots 1, 2
                if (k > 2) throw ValueError{"an inactive cell with three code-3 boundary edges is not supported"};
                t[2] = t[1]; t[1] = t[0]; t[0] = fh::kCellInactive;
            }
            s.ce0[c] = t[0]; s.ce1[c] = t[1]; s.ce2[c] = t[2];
        }
        s.area.assign(C_, 0); s.n2.assign(C_, 0); s.n1.assign(C_, 0); s.lc.assign(C_, 0); s.bed.assign(C_, 0);
        s.active.assign(C_, 1);
        for (int64_t c = 0; c < C_; ++c) {
            const real nn = (real)n[c];
            s.area[c] = (real)area[c];
            s.n2[c] = nn * nn;
            s.n1[c] = nn;
            s.lc[c] = (real)lc_[c];
            s.bed[c] = (real)bed[c];
            s.active[c] = act(c) ? 1 : 0;
        }
        s.rainm.assign(C_, real(1));
        for (int64_t c = 0; c < C_; ++c) if (!s.active[c]) s.rainm[c] = real(0);
        s.f0.assign(C_, real(0)); s.fc.assign(C_, real(0)); s.hk.assign(C_, real(0));
        s.h.assign(C_, real(0)); s.hu.assign(C_, real(0)); s.hv.assign(C_, real(0)); s.wet.assign(C_, real(0));
        s.has_surface = true;
    }

    void upload_coupling(int64_t N_, int64_t I_, int64_t S_, const int32_t* icell_, const int32_t* inode_,
                         const int32_t* nsc_, const float* inv_, const float* As_, const float* dmax_,
                         const float* capq_, const float* surq_, const int32_t* scell_, const int32_t* snode_,
                         const float* sw_, int32_t snorm_) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (N_ < 0 || I_ < 0 || S_ < 0) throw ValueError{"negative coupling sizes"};
        if (N_ > 0 && (!inv_ || !As_ || !dmax_)) throw ValueError{"node arrays must not be NULL"};
        if (I_ > 0 && (!icell_ || !inode_)) throw ValueError{"inlet arrays must not be NULL"};
        if (S_ > 0 && (!scell_ || !snode_ || !sw_)) throw ValueError{"surcharge table arrays must not be NULL"};
        for (int64_t i = 0; i < I_; ++i)
            if (icell_[i] < 0 || icell_[i] >= s.C || inode_[i] < 0 || inode_[i] >= N_)
                throw ValueError{"inlet cell/node index out of range"};
        for (int64_t n = 0; n < N_; ++n)
            if (nsc_ && nsc_[n] >= s.C) throw ValueError{"node_surface_cell out of range"};
        for (int64_t k = 0; k < S_; ++k)
            if (scell_[k] >= s.C || snode_[k] >= N_) throw ValueError{"surcharge table index out of range"};
        touch();
        s.N = N_; s.I = I_; s.S = S_;
        s.inv = conv<real>(inv_, N_); s.As = conv<real>(As_, N_); s.dmax = conv<real>(dmax_, N_);
        s.nsc = ints(nsc_, N_);
        s.icell = ints(icell_, I_); s.inode = ints(inode_, I_);
        s.capq = conv<real>(capq_, I_); s.surq = conv<real>(surq_, I_);
        s.scell = ints(scell_, S_); s.snode = ints(snode_, S_); s.sw = conv<real>(sw_, S_);
        s.snorm = snorm_ ? 1 : 0;
        s.V.assign(N_, real(0));
        if ((int64_t)s.oarea.size() != N_) {
            s.oarea.assign(N_, real(0)); s.ocoef.assign(N_, real(0)); s.otw.assign(N_, real(0)); s.ocap.assign(N_, real(0));
        }
        if (s.L > 0) {  // links refer to nodes: a new node set drops them
            s.L = 0;
            s.lfrom.clear(); s.lto.clear();
        }
    }

    void upload_pipe(int64_t L_, const int32_t* from, const int32_t* to, const float* len_, const float* n_,
                     const float* cap_, const float* inv_, const float* D_, const float* minor_,
                     const int32_t* geom_, const float* rw_, const float* rh_, const float* oa, const float* oc,
                     const float* otw_, const float* ocap_) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (L_ < 0) throw ValueError{"negative link count"};
        if (L_ > 0 && (!from || !to || !len_ || !n_ || !cap_ || !inv_ || !D_))
            throw ValueError{"link arrays must not be NULL"};
        for (int64_t l = 0; l < L_; ++l)
            if (from[l] < 0 || from[l] >= s.N || to[l] < 0 || to[l] >= s.N)
                throw ValueError{"link node index out of range (upload coupling first)"};
        touch();
        s.L = L_;
        s.lfrom = ints(from, L_); s.lto = ints(to, L_); s.geom = ints(geom_, L_, 1);
        s.llen = conv<real>(len_, L_); s.man = conv<real>(n_, L_); s.lcap = conv<real>(cap_, L_);
        s.linv = conv<real>(inv_, L_); s.D = conv<real>(D_, L_); s.minor = conv<real>(minor_, L_);
        s.rw = conv<real>(rw_, L_); s.rh = conv<real>(rh_, L_);
        s.Q.assign(L_, real(0));
        s.oarea = conv<real>(oa, s.N); s.ocoef = conv<real>(oc, s.N);
        s.otw = conv<real>(otw_, s.N); s.ocap = conv<real>(ocap_, s.N);
    }

    void upload_recharge(int64_t R_, const int32_t* cell, const float* S_, const float* cc, const float* rc,
                         const float* stop) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (R_ < 0) throw ValueError{"negative recharge row count"};
        if (R_ > 0 && (!cell || !S_ || !cc || !rc || !stop)) throw ValueError{"recharge arrays must not be NULL"};
        for (int64_t r = 0; r < R_; ++r)
            if (cell[r] < 0 || cell[r] >= s.C) throw ValueError{"recharge source_cell out of range"};
        touch();
        s.R = R_;
        s.rcell = ints(cell, R_);
        s.rS = conv<real>(S_, R_); s.rcc = conv<real>(cc, R_); s.rrc = conv<real>(rc, R_);
        s.rstop = conv<real>(stop, R_);
        s.rstore.assign(R_, real(0));
    }

    // Stage values of the code-3 edges (TriangularSurfaceForcing.boundary_stage_m,
    // triangular_backend.py:93-101): n == 1 is one stage for every edge, n == the
    // caller's edge count one per edge, n == 0 (or NULL) removes them.
    void set_boundary_stage(int64_t n, const float* stage) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (n < 0 || (n > 0 && !stage) || (n > 1 && n != s.E_caller))
            throw ValueError{"boundary_stage_m must be one value or one value per edge"};
        touch();
        if (n == 0 || !stage) s.stage.clear();
        else if (n == 1) s.stage.assign((size_t)s.E_caller, stage[0]);
        else s.stage.assign(stage, stage + n);
        const bool staged = !s.stage.empty();
        for (size_t k = 0; k < s.st_pos.size(); ++k)
            s.er[(size_t)s.st_pos[k]] = s.st_dead[k] ? (staged ? fh::kGhostDeadStage : fh::kGhostDeadOther)
                                                     : (staged ? fh::kGhostStage : fh::kGhostOther);
    }

    void set_forcing(int64_t nr, const float* rs, const float* re, const float* rr, const float* m,
                     const float* f0, const float* fc, const float* k, int64_t nt, const float* et,
                     int64_t nrcv, const int32_t* rcv, const float* q) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (nr < 0 || nt < 0 || nrcv < 0) throw ValueError{"negative forcing sizes"};
        if (nr && (!rs || !re || !rr)) throw ValueError{"rain interval arrays must not be NULL"};
        if (nrcv > 0 && (nt <= 0 || !et || !rcv || !q)) throw ValueError{"external inflow arrays incomplete"};
        for (int64_t j = 0; j < nrcv; ++j)
            if (rcv[j] < 0 || rcv[j] >= s.C) throw ValueError{"external receiver cell out of range"};
        touch();
        s.rain_s.assign(rs, rs + nr); s.rain_e.assign(re, re + nr); s.rain_r.assign(rr, rr + nr);
        for (int64_t c = 0; c < s.C; ++c) {
            // multiplier folded with the active mask (rain_depth = rate * m * dt * active)
            s.rainm[c] = s.active[c] ? (real)(m ? m[c] : 1.0f) : real(0);
            s.f0[c] = (real)(f0 ? f0[c] : 0.0f);
            s.fc[c] = (real)(fc ? fc[c] : 0.0f);
            s.hk[c] = (real)(k ? k[c] : 0.0f);
        }
        s.Rcv = nrcv;
        s.T = nrcv > 0 ? nt : 0;
        s.ext_t.assign(et ? et : nullptr, et ? et + s.T : nullptr);
        s.ext_q.assign(q ? q : nullptr, q ? q + s.T * s.Rcv : nullptr);
        s.rcv = ints(rcv, nrcv);
        // forcing-table hints start over; the forward scan is used only on sorted tables
        pull_ctl();
        fh::Ctl& c = ctl_h[cur];
        c.clk[par].rain.hint = 0;
        c.clk[par].ext_hint = 0;
        c.clk[par].rain.next_end = -INFINITY;
        c.rain_sorted = std::is_sorted(s.rain_e.begin(), s.rain_e.end()) ? 1 : 0;
        c.ext_sorted = std::is_sorted(s.ext_t.begin(), s.ext_t.end()) ? 1 : 0;
        push_ctl();
    }

    void set_state(const float* h, const float* hu, const float* hv, const float* wet, const float* vol,
                   const float* q, const float* rs) override {
        Member<real>& s = cm();
        if (!s.has_surface) throw ValueError{"fh_upload_surface must be called first"};
        if (packed && !dirty) {
            finalize_if_pending();
            // write the member's slices on the device
            const int cb = cells.base[pos_of_member(cur)];
            std::vector<V> st((size_t)s.C);
            d_state.get(st.data(), cb, s.C, stream);
            FH_CHECK(hipStreamSynchronize(stream));
            for (int64_t c = 0; c < s.C; ++c) {
                const int q = s.cp((int)c);
                if (h) st[q].x = (real)h[c];
                if (hu) st[q].y = (real)hu[c];
                if (hv) st[q].z = (real)hv[c];
            }
            d_state.put(st.data(), cb, s.C, stream);
            if (wet) {
                std::vector<real> w((size_t)s.C);
                for (int64_t c = 0; c < s.C; ++c) w[(size_t)s.cp((int)c)] = (real)wet[c];
                d_wet.put(w.data(), cb, s.C, stream);
            }
            if (P.compensated && (h || hu || hv)) {
                std::vector<V> z((size_t)s.C);
                for (auto& v : z) { v.x = 0; v.y = 0; v.z = 0; v.w = 0; }
                d_comp.put(z.data(), cb, s.C, stream);
            }
            const int mp = pos_of_member(cur);
            if (vol && s.N) {
                std::vector<real> v((size_t)s.N);
                for (int64_t n = 0; n < s.N; ++n) v[(size_t)s.np((int)n)] = (real)vol[n];
                d_V.put(v.data(), nodes.base[mp], s.N, stream);
            }
            if (q && s.L) {
                std::vector<real> v((size_t)s.L);
                for (int64_t l = 0; l < s.L; ++l) v[(size_t)s.lp((int)l)] = (real)q[l];
                d_Q.put(v.data(), links.base[mp], s.L, stream);
            }
            if (rs && s.R) { auto v = conv<real>(rs, s.R); d_rstore.put(v.data(), rrow_base[mp], s.R, stream); }
            FH_CHECK(hipStreamSynchronize(stream));
            return;
        }
        if (h) s.h = conv<real>(h, s.C);
        if (hu) s.hu = conv<real>(hu, s.C);
        if (hv) s.hv = conv<real>(hv, s.C);
        if (wet) s.wet = conv<real>(wet, s.C);
        if (vol && s.N) s.V = conv<real>(vol, s.N);
        if (q && s.L) s.Q = conv<real>(q, s.L);
        if (rs && s.R) s.rstore = conv<real>(rs, s.R);
    }

    void get_state(float* h, float* hu, float* hv, float* wet, float* vol, float* q, float* rs) override {
        ensure_packed();
        finalize_if_pending();
        Member<real>& s = cm();
        const int mp = pos_of_member(cur);
        const int cb = cells.base[mp];
        std::vector<V> st((size_t)s.C);
        std::vector<real> w((size_t)s.C), v((size_t)s.N), qq((size_t)s.L), r((size_t)s.R);
        d_state.get(st.data(), cb, s.C, stream);
        d_wet.get(w.data(), cb, s.C, stream);
        if (s.N) d_V.get(v.data(), nodes.base[mp], s.N, stream);
        if (s.L) d_Q.get(qq.data(), links.base[mp], s.L, stream);
        if (s.R) d_rstore.get(r.data(), rrow_base[mp], s.R, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        for (int64_t c = 0; c < s.C; ++c) {
            const int q = s.cp((int)c);
            if (h) h[c] = (float)st[q].x;
            if (hu) hu[c] = (float)st[q].y;
            if (hv) hv[c] = (float)st[q].z;
            if (wet) wet[c] = (float)w[q];
        }
        if (vol) for (int64_t n = 0; n < s.N; ++n) vol[n] = (float)v[(size_t)s.np((int)n)];
        if (q) for (int64_t l = 0; l < s.L; ++l) q[l] = (float)qq[(size_t)s.lp((int)l)];
        if (rs) for (int64_t k = 0; k < s.R; ++k) rs[k] = (float)r[k];
    }

    void get_state_f64(double* h, double* hu, double* hv) override {
        ensure_packed();
        finalize_if_pending();
        Member<real>& s = cm();
        std::vector<V> st((size_t)s.C);
        d_state.get(st.data(), cells.base[pos_of_member(cur)], s.C, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        for (int64_t c = 0; c < s.C; ++c) {
            const int q = s.cp((int)c);
            if (h) h[c] = (double)st[q].x;
            if (hu) hu[c] = (double)st[q].y;
            if (hv) hv[c] = (double)st[q].z;
        }
    }

    void get_drainage_f64(double* vol, double* q) override {
        ensure_packed();
        Member<real>& s = cm();
        const int mp = pos_of_member(cur);
        std::vector<real> v((size_t)s.N), qq((size_t)s.L);
        if (s.N) d_V.get(v.data(), nodes.base[mp], s.N, stream);
        if (s.L) d_Q.get(qq.data(), links.base[mp], s.L, stream);
        FH_CHECK(hipStreamSynchronize(stream));
        if (vol) for (int64_t n = 0; n < s.N; ++n) vol[n] = (double)v[(size_t)s.np((int)n)];
        if (q) for (int64_t l = 0; l < s.L; ++l) q[l] = (double)qq[(size_t)s.lp((int)l)];
    }

    // ------------------------------------------------------------------ control blocks
    void pull_ctl() {
        FH_CHECK(hipMemcpyAsync(ctl_h, ctl, sizeof(fh::Ctl) * M, hipMemcpyDeviceToHost, stream));
        FH_CHECK(hipStreamSynchronize(stream));
        ctl_fresh = true;
    }
    void push_ctl() {
        FH_CHECK(hipMemcpyAsync(ctl, ctl_h, sizeof(fh::Ctl) * M, hipMemcpyHostToDevice, stream));
        // no step is in flight: every CFL minimum slot starts the next step at +inf
        FH_CHECK(hipMemcpyAsync(d_cfl.p, cfl_inf.data(), cfl_inf.size() * sizeof(unsigned long long),
                                hipMemcpyHostToDevice, stream));
        FH_CHECK(hipStreamSynchronize(stream));
        ctl_fresh = true;
    }
    // ctl_h is indexed by member id (not pack position)
    void set_clock(double t, double cached, int64_t step) override {
        pull_ctl();
        fh::Clock& c = ctl_h[cur].clk[par];
        c.t = t;
        c.cached_cfl = (double)(real)cached;
        c.step = step;
        c.rain.hint = 0;
        c.ext_hint = 0;
        c.rain.next_end = -INFINITY;
        c.eval = (P.fixed_dt <= 0.0 && (step % kp.interval) == 0) ? 1 : 0;
        push_ctl();
    }
    void get_clock(fh_clock* c) override {
        if (!ctl_fresh) pull_ctl();
        const fh::Clock& x = ctl_h[cur].clk[par];
        c->time_s = x.t;
        c->cached_dt_s = x.cached_cfl;
        c->last_dt_s = ctl_h[cur].dt;
        c->global_step = x.step;
        c->interval_steps = x.interval_step;
        c->cfl_evaluations = x.cfl_evals;
    }
    void set_ledger(const double* l) override {
        pull_ctl();
        for (int k = 0; k < fh::kLedgerTerms; ++k) ctl_h[cur].ledger[k] = l ? l[k] : 0.0;
        push_ctl();
    }
    void get_ledger(double* l) override {
        if (!ctl_fresh) pull_ctl();
        for (int k = 0; k < fh::kLedgerTerms; ++k) l[k] = ctl_h[cur].ledger[k];
    }

    // ------------------------------------------------------------------ packing
    int pos_of_member(int m) const { return member_pos[(size_t)m]; }

    void pull_all() {
        finalize_if_pending();
        // copy device state back into every member's host staging
        for (int m = 0; m < M; ++m) {
            Member<real>& s = mem[(size_t)m];
            if (!s.has_surface) continue;
            const int mp = pos_of_member(m);
            std::vector<V> st((size_t)s.C);
            std::vector<real> w((size_t)s.C);
            d_state.get(st.data(), cells.base[mp], s.C, stream);
            d_wet.get(w.data(), cells.base[mp], s.C, stream);
            std::vector<real> v((size_t)s.N), q((size_t)s.L);
            const bool hv_ = s.N && (int64_t)s.V.size() == s.N, hq = s.L && (int64_t)s.Q.size() == s.L;
            if (hv_) d_V.get(v.data(), nodes.base[mp], s.N, stream);
            if (hq) d_Q.get(q.data(), links.base[mp], s.L, stream);
            if (s.R && (int64_t)s.rstore.size() == s.R) d_rstore.get(s.rstore.data(), rrow_base[mp], s.R, stream);
            FH_CHECK(hipStreamSynchronize(stream));
            for (int64_t c = 0; c < s.C; ++c) {
                const int p = s.cp((int)c);
                s.h[c] = st[p].x; s.hu[c] = st[p].y; s.hv[c] = st[p].z; s.wet[c] = w[p];
            }
            if (hv_) for (int64_t n = 0; n < s.N; ++n) s.V[n] = v[(size_t)s.np((int)n)];
            if (hq) for (int64_t l = 0; l < s.L; ++l) s.Q[l] = q[(size_t)s.lp((int)l)];
        }
    }

    void ensure_packed() {
        for (int m = 0; m < M; ++m)
            if (!mem[(size_t)m].has_surface) throw ValueError{"fh_upload_surface must be called for every member"};
        if (!packed || dirty) pack();
    }

    void pack() {
        drop_graph();
        ctl_fresh = false;
        wet_pending = false;
        for (size_t i = 0; i < mem.size(); ++i) {
            Member<real>& m = mem[i];
            if (P.cell_order == 1 && m.C > 0) {
                if (m.bfs.empty()) {
                    // a member with an earlier member's mesh takes its order
                    for (size_t j = 0; j < i && m.bfs.empty(); ++j)
                        if (!mem[j].bfs.empty() && same_surface(mem[j], m)) m.bfs = mem[j].bfs;
                    if (m.bfs.empty()) locality_order(m);
                }
                assign_positions(m);
            } else {
                m.cpos.clear(); m.epos.clear(); m.npos.clear(); m.lpos.clear();
            }
        }
        // A batch whose members all have the same mesh (same device layout) keeps
        // one copy of the edge tables and the cell->edge table, with
        // member-local indices; its members keep their given order.
        shared = P.share_mesh != 0 && M > 1 && same_mesh();
        if (shared) {
            // exact block -> member division (shared_block): blocks * blocks per member < 2^40,
            // and block index * reciprocal (about M * 2^40) within 64 bits
            const long long cb = (mem[0].C + bs - 1) / bs, eb = (mem[0].E + (int64_t)bs * ept - 1) / ((int64_t)bs * ept);
            const long long mx = std::max(cb, eb);
            if (cb == 0 || eb == 0 || M >= (1 << 20) || (long double)M * mx * mx >= (long double)(1ull << 40))
                shared = false;
        }
        // member order in the concatenation
        order.resize(M);
        std::iota(order.begin(), order.end(), 0);
        if (shared) {
            // one size: the given order
        } else if (P.pack_order == 1) {
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return mem[a].C > mem[b].C; });
        } else if (P.pack_order == 2) {   // alternate largest / smallest
            std::vector<int> bysize(order);
            std::stable_sort(bysize.begin(), bysize.end(), [&](int a, int b) { return mem[a].C > mem[b].C; });
            int lo = 0, hi = M - 1, k = 0;
            while (lo <= hi) { order[k++] = bysize[lo++]; if (lo <= hi) order[k++] = bysize[hi--]; }
        }
        member_pos.assign(M, 0);
        for (int i = 0; i < M; ++i) member_pos[order[i]] = i;
        auto counts = [&](auto fn) {
            std::vector<int64_t> v(M);
            for (int i = 0; i < M; ++i) v[i] = fn(mem[order[i]]);
            return v;
        };
        cells.build(counts([](const Member<real>& s) { return s.C; }), bs);
        edges.build(counts([](const Member<real>& s) { return s.E; }), bs * ept);
        nodes.build(counts([](const Member<real>& s) { return s.N; }), fh::kNodeBlock);
        links.build(counts([](const Member<real>& s) { return s.L; }), 256);
        // per-member derived lists (inlet cells, return cells, receiver cells)
        std::vector<std::vector<int>> icells(M), rcells(M), xcells(M);
        for (int i = 0; i < M; ++i) {
            const Member<real>& s = mem[order[i]];
            std::vector<char> mark((size_t)s.C, 0);
            for (int64_t k = 0; k < s.I; ++k) mark[s.cp(s.icell[k])] = 1;
            for (int64_t c = 0; c < s.C; ++c) if (mark[c]) icells[i].push_back((int)c);
            std::fill(mark.begin(), mark.end(), 0);
            for (int64_t k = 0; k < s.S; ++k) if (s.snode[k] >= 0 && s.scell[k] >= 0) mark[s.cp(s.scell[k])] = 1;
            std::vector<int> expl((size_t)s.N, 0);
            for (int64_t k = 0; k < s.S; ++k) if (s.snode[k] >= 0 && s.scell[k] >= 0) expl[s.snode[k]] = 1;
            for (int64_t n = 0; n < s.N; ++n) if (!expl[n] && s.nsc[n] >= 0) mark[s.cp(s.nsc[n])] = 1;
            for (int64_t r = 0; r < s.R; ++r) mark[s.cp(s.rcell[r])] = 1;
            for (int64_t c = 0; c < s.C; ++c) if (mark[c]) rcells[i].push_back((int)c);
            std::fill(mark.begin(), mark.end(), 0);
            for (int64_t j = 0; j < s.Rcv; ++j) mark[s.cp(s.rcv[j])] = 1;
            for (int64_t c = 0; c < s.C; ++c) if (mark[c]) xcells[i].push_back((int)c);
        }
        icl.build(counts_of(icells), fh::kInletRows);
        rcl.build(counts_of(rcells), 256);
        xcl.build(counts_of(xcells), 256);
        // k_cfl / k_li_cfl blocks: one per 2,048 of the member's cells (at most 512),
        // each thread looping over 8 cells; the block minimums spread over the CFL slots
        {
            std::vector<int64_t> nb(M);
            for (int i = 0; i < M; ++i) nb[i] = std::max<int64_t>(1, std::min<int64_t>(512, (mem[order[i]].C + 2047) / 2048));
            begin_l.base.assign(M, 0); begin_l.count.assign(M, 0);
            begin_l.blk_member.clear(); begin_l.mem_blk0.assign(M, 0); begin_l.mem_nblk.assign(M, 0);
            begin_l.mem_end.assign(M, 0);
            int b = 0;
            for (int i = 0; i < M; ++i) {
                begin_l.mem_blk0[i] = b;
                begin_l.mem_nblk[i] = (int)nb[i];
                for (int k = 0; k < nb[i]; ++k) begin_l.blk_member.push_back(i);
                b += (int)nb[i];
            }
            begin_l.blocks = b;
            begin_l.tile = 256;
        }
        const hipStream_t s = stream;
        // ---------------- surface arrays
        // edge tables and the cell->edge table: one member's worth when shared
        const int sh_e = shared ? edges.total / M : edges.total, sh_c = shared ? cells.total / M : cells.total;
        std::vector<int> el(sh_e, 0), er(sh_e, fh::kGhostCopy);
        std::vector<real> nx(sh_e, 0), ny(sh_e, 0), len(sh_e, 0);
        std::vector<int> ce0(sh_c, fh::kCellInactive), ce1(sh_c, -1), ce2(sh_c, -1);
        std::vector<V> st(cells.total);
        std::vector<real> wet(cells.total, 0), area(cells.total, 1), n2(cells.total, 0), lc(cells.total, 0),
            rainm(cells.total, 0), f0(cells.total, 0), fc(cells.total, 0), hk(cells.total, 0);
        for (auto& v : st) { v.x = 0; v.y = 0; v.z = 0; v.w = 0; }
        std::vector<int> cell0(M), cell1(M), rain_off(M), rain_n(M), ext_off(M), ext_n(M);
        std::vector<float> rs_all, re_all, rr_all, et_all;
        for (int i = 0; i < M; ++i) {
            const Member<real>& m = mem[order[i]];
            const int cb = cells.base[i], eb = edges.base[i];
            // offsets of the tables' cell and edge indices (0: member-local, shared)
            const int tcb = shared ? 0 : cb, teb = shared ? 0 : eb;
            const bool fill = !shared || i == 0;
            const bool perm = !m.epos.empty();
            for (int64_t e = 0; fill && e < m.E; ++e) {
                const int64_t q = teb + (perm ? m.epos[(size_t)e] : e);
                el[q] = m.cp(m.el[e]) + tcb;
                er[q] = m.er[e] >= 0 ? m.cp(m.er[e]) + tcb : m.er[e];
                nx[q] = m.nx[e]; ny[q] = m.ny[e]; len[q] = m.len[e];
            }
            // entry = (edge position << 2) | (boundary edge << 1) | right side
            auto enc = [&](int x) {
                if (x < 0) return x;
                const int staged = x >> 1;
                const int p = perm ? m.epos[(size_t)staged] : staged;
                const int ghost = staged >= m.E_int ? 1 : 0;
                return ((p + teb) << 2) | (ghost << 1) | (x & 1);
            };
            for (int64_t c = 0; c < m.C; ++c) {
                const int64_t q = cb + m.cp((int)c);
                if (fill) {
                    const int64_t qt = tcb + m.cp((int)c);
                    ce0[qt] = enc(m.ce0[c]); ce1[qt] = enc(m.ce1[c]); ce2[qt] = enc(m.ce2[c]);
                }
                st[q].x = m.h[c]; st[q].y = m.hu[c]; st[q].z = m.hv[c]; st[q].w = m.bed[c];
                wet[q] = m.wet[c]; area[q] = m.area[c]; n2[q] = m.n2[c]; lc[q] = m.lc[c];
                rainm[q] = m.rainm[c]; f0[q] = m.f0[c]; fc[q] = m.fc[c]; hk[q] = m.hk[c];
            }
            cell0[i] = cb; cell1[i] = cb + (int)m.C;
            rain_off[i] = (int)rs_all.size(); rain_n[i] = (int)m.rain_s.size();
            rs_all.insert(rs_all.end(), m.rain_s.begin(), m.rain_s.end());
            re_all.insert(re_all.end(), m.rain_e.begin(), m.rain_e.end());
            rr_all.insert(rr_all.end(), m.rain_r.begin(), m.rain_r.end());
            ext_off[i] = (int)et_all.size(); ext_n[i] = m.Rcv > 0 ? (int)m.T : 0;
            if (m.Rcv > 0) et_all.insert(et_all.end(), m.ext_t.begin(), m.ext_t.end());
        }
        d_el.set(el, s); d_er.set(er, s); d_nx.set(nx, s); d_ny.set(ny, s); d_len.set(len, s);
        // stage ghost depth max(stage - z_left, 0) of every code-3 edge of the members
        // with stage values, in the state precision, at the edge's device position
        {
            bool any = false;
            for (const auto& m : mem) any = any || !m.stage.empty();
            std::vector<real> gst(any ? (size_t)edges.total : 0, real(0));
            for (int i = 0; any && i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                for (size_t k = 0; k < m.stage.size() && k < m.st_pos.size(); ++k) {
                    const int e = m.st_pos[k];
                    const real d = (real)m.stage[(size_t)m.st_edge[k]] - m.bed[(size_t)m.el[(size_t)e]];
                    const int64_t q = edges.base[i] + (m.epos.empty() ? e : m.epos[(size_t)e]);
                    gst[(size_t)q] = (d != d || d > real(0)) ? d : real(0);
                }
            }
            d_stage_h.set(gst, s);
        }
        d_rec.alloc(std::max(edges.total, 1)); d_recR.alloc(std::max(edges.total, 1));
        d_ce0.set(ce0, s); d_ce1.set(ce1, s); d_ce2.set(ce2, s);
        if (li) {
            // Manning n, and each cell's interior entry slots in ascending edge order
            // (two bits per slot, 3 = end) for the donor sums
            std::vector<real> man(cells.total, 0);
            std::vector<unsigned char> perm(sh_c, 0xFF);
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                const int cb = cells.base[i];
                for (int64_t c = 0; c < m.C; ++c) {
                    const int64_t q = cb + m.cp((int)c);
                    man[q] = m.n1[c];
                    if (shared && i > 0) continue;
                    const int t[3] = {m.ce0[c], m.ce1[c], m.ce2[c]};
                    int slot[3], key[3], n = 0;
                    for (int k = 0; k < 3; ++k)
                        if (t[k] >= 0 && (t[k] >> 1) < m.E_int) { slot[n] = k; key[n] = t[k] >> 1; ++n; }
                    for (int a = 1; a < n; ++a)
                        for (int b2 = a; b2 > 0 && key[b2 - 1] > key[b2]; --b2) {
                            std::swap(key[b2 - 1], key[b2]);
                            std::swap(slot[b2 - 1], slot[b2]);
                        }
                    unsigned code = 0xFF;
                    for (int k = 0; k < n; ++k) code = (code & ~(3u << (2 * k))) | ((unsigned)slot[k] << (2 * k));
                    perm[(shared ? 0 : cb) + m.cp((int)c)] = (unsigned char)code;
                }
            }
            d_cell_n.set(man, s);
            d_perm.set(perm, s);
            d_qli.alloc(std::max(edges.total, 1));
        } else {
            d_cell_n.release();
            d_perm.release();
            d_qli.release();
        }
        d_state.set(st, s); d_wet.set(wet, s); d_area.set(area, s); d_n2.set(n2, s); d_lc.set(lc, s);
        d_rainm.set(rainm, s); d_f0.set(f0, s); d_fc.set(fc, s); d_k.set(hk, s);
        build_source_classes(n2, rainm, f0, fc, hk);
        if (P.compensated) {
            std::vector<V> z(cells.total);
            for (auto& v : z) { v.x = 0; v.y = 0; v.z = 0; v.w = 0; }
            d_comp.set(z, s);
        }
        d_rain_s.set(rs_all, s); d_rain_e.set(re_all, s); d_rain_r.set(rr_all, s); d_ext_t.set(et_all, s);
        seg_cells.set(cells, order, s); seg_edges.set(edges, order, s); seg_begin.set(begin_l, order, s);
        d_cell0.set(by_id(cell0, order), s); d_cell1.set(by_id(cell1, order), s);
        d_rain_off.set(by_id(rain_off, order), s); d_rain_n.set(by_id(rain_n, order), s);
        d_ext_off.set(by_id(ext_off, order), s); d_ext_n.set(by_id(ext_n, order), s);
        d_part_upd.set(std::vector<double>(4 * (size_t)std::max(cells.blocks, 1), 0.0), s);
        // cells no kept edge reaches: the flux kernel's CFL evaluation covers them in extra blocks
        {
            std::vector<std::vector<int>> orph(M);
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                std::vector<char> seen((size_t)m.C, 0);
                for (int64_t e = 0; e < m.E; ++e) {
                    seen[(size_t)m.cp(m.el[e])] = 1;
                    if (m.er[e] >= 0) seen[(size_t)m.cp(m.er[e])] = 1;
                }
                for (int64_t c = 0; c < m.C; ++c) if (!seen[(size_t)c]) orph[i].push_back((int)c + cells.base[i]);
            }
            orphans.build(counts_of(orph), 256);
            std::vector<int> oc((size_t)orphans.total, 0);
            for (int i = 0; i < M; ++i)
                for (size_t k = 0; k < orph[i].size(); ++k) oc[(size_t)orphans.base[i] + k] = orph[i][k];
            d_ocells.set(oc, s);
            seg_orph.set(orphans, order, s);
        }
        pack_drainage(icells, rcells, xcells, area);
        // control blocks: per member, unchanged
        FH_CHECK(hipStreamSynchronize(s));
        packed = true;
        dirty = false;
        any_drainage = nodes.total > 0;
    }

    template <class A>
    static bool same_array(const A& x, const A& y) {
        return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(x[0])) == 0);
    }
    // The staged surfaces of a and b are the same mesh: sizes, edge tables and
    // cell->edge entries equal bit for bit.
    static bool same_surface(const Member<real>& a, const Member<real>& b) {
        return b.C == a.C && b.E == a.E && b.E_int == a.E_int && same_array(a.el, b.el) && same_array(a.er, b.er) &&
               same_array(a.nx, b.nx) && same_array(a.ny, b.ny) && same_array(a.len, b.len) &&
               same_array(a.ce0, b.ce0) && same_array(a.ce1, b.ce1) && same_array(a.ce2, b.ce2);
    }
    // Every member has member 0's mesh in the same device layout: the tables a
    // shared-mesh batch stores once (edges, cell->edge entries) would be equal.
    bool same_mesh() const {
        for (int i = 1; i < M; ++i) {
            const Member<real>& b = mem[(size_t)i];
            if (!same_surface(mem[0], b) || !same_array(mem[0].cpos, b.cpos) || !same_array(mem[0].epos, b.epos))
                return false;
        }
        return true;
    }

    // Cell and edge index spaces as the surface kernels see them (with the
    // shared-mesh strides when the batch shares one mesh).
    fh::Seg cell_seg() const {
        fh::Seg g = seg_cells.seg();
        if (shared) set_shared(g, cells.mem_nblk[0], cells.total / M, edges.total / M);
        return g;
    }
    fh::Seg edge_seg() const {
        fh::Seg g = seg_edges.seg();
        if (shared) set_shared(g, edges.mem_nblk[0], edges.total / M, cells.total / M);
        return g;
    }
    void set_shared(fh::Seg& g, int nblk, int items, int link) const {
        g.sh_members = M; g.sh_nblk = nblk; g.sh_items = items; g.sh_link = link;
        g.sh_rot = nblk / M;
        g.sh_recip = (1ull << 40) / (unsigned long long)nblk + 1;
    }

    // Class table of the per-cell source tuples (n^2, m, f0, fc, k), compared bit
    // for bit; used by the update when the batch has at most 65,535 of them.
    void build_source_classes(const std::vector<real>& n2, const std::vector<real>& rm, const std::vector<real>& f0,
                              const std::vector<real>& fc, const std::vector<real>& hk) {
        struct Key {
            real v[5];
            bool operator<(const Key& o) const { return std::memcmp(v, o.v, sizeof(v)) < 0; }
        };
        std::map<Key, int> ids;
        std::vector<unsigned short> cls(n2.size(), 0);
        std::vector<V> tab;
        std::vector<real> tab_hk;
        use_classes = true;
        for (size_t c = 0; c < n2.size(); ++c) {
            Key k{{n2[c], rm[c], f0[c], fc[c], hk[c]}};
            auto it = ids.find(k);
            if (it == ids.end()) {
                if (ids.size() >= 65535) { use_classes = false; break; }
                it = ids.emplace(k, (int)tab.size()).first;
                V t;
                t.x = n2[c]; t.y = rm[c]; t.z = f0[c]; t.w = fc[c];
                tab.push_back(t);
                tab_hk.push_back(hk[c]);
            }
            cls[c] = (unsigned short)it->second;
        }
        if (!P.source_classes) use_classes = false;
        if (use_classes) {
            d_cls.set(cls, stream); d_cls_tab.set(tab, stream); d_cls_hk.set(tab_hk, stream);
        } else {
            d_cls.release(); d_cls_tab.release(); d_cls_hk.release();
        }
    }

    static std::vector<int64_t> counts_of(const std::vector<std::vector<int>>& v) {
        std::vector<int64_t> c(v.size());
        for (size_t i = 0; i < v.size(); ++i) c[i] = (int64_t)v[i].size();
        return c;
    }

    void pack_drainage(const std::vector<std::vector<int>>& icells, const std::vector<std::vector<int>>& rcells,
                       const std::vector<std::vector<int>>& xcells, const std::vector<real>& area_of_cell) {
        const hipStream_t s = stream;
        const int NT = nodes.total, LT = links.total;
        // global item bases for inlets, entries, recharge rows, receivers (no padding needed)
        std::vector<int> ibase(M), sbase(M), rbase(M), xbase(M), qbase(M);
        int It = 0, St = 0, Rt = 0, Xt = 0, Qt = 0;
        for (int i = 0; i < M; ++i) {
            const Member<real>& m = mem[order[i]];
            ibase[i] = It; It += (int)m.I;
            sbase[i] = St; St += (int)m.S;
            rbase[i] = Rt; Rt += (int)m.R;
            xbase[i] = Xt; Xt += (int)m.Rcv;
            qbase[i] = Qt; Qt += (int)(m.T * m.Rcv);
        }
        rrow_base = rbase;
        std::vector<real> inv(NT, 0), As(NT, 1), dmax(NT, 0), oa(NT, 0), oc(NT, 0), otw(NT, 0), ocap(NT, 0), V(NT, 0);
        std::vector<int> nsc(NT, -1), expl(NT, 0);
        std::vector<int> ni_ptr(NT + 1, 0), nl_ptr(NT + 1, 0), ne_ptr(NT + 1, 0);
        std::vector<int> ni_idx, nl_ent, ne_idx;
        std::vector<int> icell(std::max(It, 0)), inode(std::max(It, 0));
        std::vector<real> capq(It), surq(It);
        std::vector<int> from(LT, 0), to(LT, 0), geom(LT, 1);
        std::vector<real> llen(LT, 1), man(LT, 0), lcap(LT, 0), linv(LT, 0), D(LT, 1), minor(LT, 0), rw(LT, 0),
            rh(LT, 0), Q(LT, 0);
        std::vector<int> snode(St, -1);
        std::vector<real> sw(St, 0);
        std::vector<int> snorm(M, 0);
        std::vector<real> rS(Rt), rcc(Rt), rrc(Rt), rstop(Rt), rstore(Rt);
        std::vector<float> q_all;
        std::vector<int> rcv_qbase(Xt), rcv_stride(Xt);
        std::vector<int> ic_cells(icl.total, 0), ic_ptr(icl.total + 1, 0), ic_idx;
        std::vector<int> rc_cells(rcl.total, 0), rc_eptr(rcl.total + 1, 0), rc_eidx, rc_fptr(rcl.total + 1, 0), rc_fidx,
            rc_rptr(rcl.total + 1, 0), rc_ridx;
        std::vector<int> xc_cells(xcl.total, 0), xc_ptr(xcl.total + 1, 0), xc_idx;
        for (int i = 0; i < M; ++i) {
            const Member<real>& m = mem[order[i]];
            const int cb = cells.base[i], nb = nodes.base[i], lb = links.base[i];
            snorm[i] = m.snorm;
            for (int64_t n = 0; n < m.N; ++n) {
                const int g = nb + m.np((int)n);
                inv[g] = m.inv[n]; As[g] = m.As[n]; dmax[g] = m.dmax[n];
                oa[g] = m.oarea[n]; oc[g] = m.ocoef[n]; otw[g] = m.otw[n]; ocap[g] = m.ocap[n];
                V[g] = m.V[n];
                nsc[g] = m.nsc[n] >= 0 ? m.cp(m.nsc[n]) + cb : -1;
            }
            for (int64_t k = 0; k < m.I; ++k) {
                icell[ibase[i] + k] = m.cp(m.icell[k]) + cb;
                inode[ibase[i] + k] = m.np(m.inode[k]) + nb;
                capq[ibase[i] + k] = m.capq[k]; surq[ibase[i] + k] = m.surq[k];
            }
            for (int64_t l = 0; l < m.L; ++l) {
                const int g = lb + m.lp((int)l);
                from[g] = m.np(m.lfrom[l]) + nb; to[g] = m.np(m.lto[l]) + nb; geom[g] = m.geom[l];
                llen[g] = m.llen[l]; man[g] = m.man[l]; lcap[g] = m.lcap[l]; linv[g] = m.linv[l];
                D[g] = m.D[l]; minor[g] = m.minor[l]; rw[g] = m.rw[l]; rh[g] = m.rh[l];
                Q[g] = m.Q[l];
            }
            for (int64_t k = 0; k < m.S; ++k) {
                const bool valid = m.snode[k] >= 0 && m.scell[k] >= 0;
                snode[sbase[i] + k] = valid ? m.np(m.snode[k]) + nb : -1;
                sw[sbase[i] + k] = m.sw[k];
            }
            for (int64_t r = 0; r < m.R; ++r) {
                rS[rbase[i] + r] = m.rS[r]; rcc[rbase[i] + r] = m.rcc[r]; rrc[rbase[i] + r] = m.rrc[r];
                rstop[rbase[i] + r] = m.rstop[r]; rstore[rbase[i] + r] = m.rstore[r];
            }
            for (int64_t j = 0; j < m.Rcv; ++j) {
                rcv_qbase[xbase[i] + j] = qbase[i] + (int)j;
                rcv_stride[xbase[i] + j] = (int)m.Rcv;
            }
            q_all.insert(q_all.end(), m.ext_q.begin(), m.ext_q.end());
        }
        // node CSRs (global): node -> inlets, node -> links (ascending link, orientation bit), node -> entries
        {
            std::vector<int> rows(It);
            for (int k = 0; k < It; ++k) rows[k] = inode[k];
            build_csr(rows, NT, ni_ptr, ni_idx);
            std::vector<int> cnt(NT + 1, 0);
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                for (int64_t l = 0; l < m.L; ++l) {
                    const int g = links.base[i] + m.lp((int)l);
                    cnt[from[g] + 1]++;
                    cnt[to[g] + 1]++;
                }
            }
            for (int n = 0; n < NT; ++n) cnt[n + 1] += cnt[n];
            nl_ent.assign(cnt[NT], 0);
            std::vector<int> fill(cnt.begin(), cnt.end() - 1);
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                for (int64_t l = 0; l < m.L; ++l) {   // the caller's link order (reference scatter order)
                    const int g = links.base[i] + m.lp((int)l);
                    nl_ent[fill[from[g]]++] = g << 1;
                    nl_ent[fill[to[g]]++] = (g << 1) | 1;
                }
            }
            nl_ptr = cnt;
            build_csr(snode, NT, ne_ptr, ne_idx);
            for (int n = 0; n < NT; ++n) expl[n] = ne_ptr[n + 1] > ne_ptr[n] ? 1 : 0;
        }
        // inlet cells -> inlets
        for (int i = 0; i < M; ++i) {
            const Member<real>& m = mem[order[i]];
            std::vector<int> slot((size_t)m.C, -1);
            for (size_t j = 0; j < icells[i].size(); ++j) {
                slot[icells[i][j]] = icl.base[i] + (int)j;
                ic_cells[icl.base[i] + j] = icells[i][j] + cells.base[i];
            }
            std::vector<int> rows(It, -1);
            (void)rows;
            for (int64_t k = 0; k < m.I; ++k) ic_ptr[slot[m.cp(m.icell[k])] + 1]++;
        }
        for (int j = 0; j < icl.total; ++j) ic_ptr[j + 1] += ic_ptr[j];
        ic_idx.assign(ic_ptr[icl.total], 0);
        {
            std::vector<int> fill(ic_ptr.begin(), ic_ptr.end() - 1);
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                std::vector<int> slot((size_t)m.C, -1);
                for (size_t j = 0; j < icells[i].size(); ++j) slot[icells[i][j]] = icl.base[i] + (int)j;
                for (int64_t k = 0; k < m.I; ++k) ic_idx[fill[slot[m.cp(m.icell[k])]]++] = ibase[i] + (int)k;
            }
        }
        // return cells -> entries, fallback nodes, recharge rows; receiver cells -> receivers
        {
            std::vector<int> erow(St, -1), rrow(Rt, -1), xrow(Xt, -1);
            // fallback rows in the caller's node order (the reference's scatter order), with the node
            std::vector<int> frow, fnode;
            for (int i = 0; i < M; ++i) {
                const Member<real>& m = mem[order[i]];
                std::vector<int> slot((size_t)m.C, -1), xslot((size_t)m.C, -1);
                for (size_t j = 0; j < rcells[i].size(); ++j) {
                    slot[rcells[i][j]] = rcl.base[i] + (int)j;
                    rc_cells[rcl.base[i] + j] = rcells[i][j] + cells.base[i];
                }
                for (size_t j = 0; j < xcells[i].size(); ++j) {
                    xslot[xcells[i][j]] = xcl.base[i] + (int)j;
                    xc_cells[xcl.base[i] + j] = xcells[i][j] + cells.base[i];
                }
                for (int64_t k = 0; k < m.S; ++k)
                    if (m.snode[k] >= 0 && m.scell[k] >= 0) erow[sbase[i] + k] = slot[m.cp(m.scell[k])];
                for (int64_t n = 0; n < m.N; ++n) {
                    const int g = nodes.base[i] + m.np((int)n);
                    frow.push_back(!expl[g] && m.nsc[n] >= 0 ? slot[m.cp(m.nsc[n])] : -1);
                    fnode.push_back(g);
                }
                for (int64_t r = 0; r < m.R; ++r) rrow[rbase[i] + r] = slot[m.cp(m.rcell[r])];
                for (int64_t j = 0; j < m.Rcv; ++j) xrow[xbase[i] + j] = xslot[m.cp(m.rcv[j])];
            }
            build_csr(erow, rcl.total, rc_eptr, rc_eidx);
            build_csr(frow, rcl.total, rc_fptr, rc_fidx);
            for (auto& k : rc_fidx) k = fnode[(size_t)k];
            build_csr(rrow, rcl.total, rc_rptr, rc_ridx);
            build_csr(xrow, xcl.total, xc_ptr, xc_idx);
        }
        d_inv.set(inv, s); d_As.set(As, s); d_dmax.set(dmax, s); d_V.set(V, s);
        {
            // node flags and the outfall slots: a node whose coefficient x area is
            // zero (with finite tailwater and capacity) discharges exactly zero
            std::vector<int> nflag(NT, 0);
            typedef typename fh::V4<real>::type V4r;
            std::vector<V4r> otab;
            for (int n = 0; n < NT; ++n) {
                const real prod = oc[n] * oa[n];
                const bool none = prod == real(0) && std::isfinite(otw[n]) && std::isfinite(ocap[n]);
                int f = (!expl[n] && nsc[n] >= 0) ? 1 : 0;
                if (!none) {
                    V4r o;
                    o.x = otw[n]; o.y = oc[n]; o.z = oa[n]; o.w = ocap[n];
                    otab.push_back(o);
                    f |= (int)otab.size() << 1;
                }
                nflag[n] = f;
            }
            d_nflag.set(nflag, s);
            d_outfall.set(otab, s);
        }
        std::vector<real> zN(NT, 0);
        d_Vp.set(zN, s);
        d_sfac.set(zN, s); d_over.set(zN, s); d_fb.set(zN, s); d_wsum.set(zN, s);
        // composed index maps: each kernel reads per-slot data in the order it consumes it
        {
            const int nslots = (int)ni_idx.size();
            std::vector<int> ni_cell(nslots), slot_of_inlet(std::max(It, 0), -1);
            std::vector<real> ni_surq(nslots);
            for (int k = 0; k < nslots; ++k) {
                const int i = ni_idx[k];
                ni_cell[k] = icell[i];
                ni_surq[k] = surq[i];
                slot_of_inlet[i] = k;
            }
            const int nic = (int)ic_idx.size();
            std::vector<int> ic_node(nic), ic_slot(nic);
            std::vector<real> ic_capq(nic);
            for (int k = 0; k < nic; ++k) {
                const int i = ic_idx[k];
                ic_node[k] = inode[i];
                ic_capq[k] = capq[i];
                ic_slot[k] = slot_of_inlet[i];
            }
            // link work items per block of kNodeBlock nodes: the links leaving a node of
            // the block (at its 'from' entry), and the links entering it from another
            // block (at its 'to' entry); each entry reads its link's item
            const int NB = fh::kNodeBlock;
            if (nodes.tile != NB) throw ValueError{"node layout tile differs from the node kernels' block"};
            const int nblk = (NT + NB - 1) / NB;
            std::vector<int> lb_ptr(nblk + 1, 0), it_l, it_a, it_b, nl_src(nl_ent.size(), 0);
            std::vector<int> from_item((size_t)std::max(LT, 1), -1);   // block-local item of each link's 'from' entry
            size_t max_items = 1;
            for (int bk = 0; bk < nblk; ++bk) {
                lb_ptr[bk] = (int)it_l.size();
                const int n0 = bk * NB, n1 = std::min(NT, n0 + NB);
                for (int n = n0; n < n1; ++n)
                    for (int k = nl_ptr[n]; k < nl_ptr[n + 1]; ++k) {
                        const int l = nl_ent[k] >> 1, is_to = nl_ent[k] & 1;
                        if (!is_to || from[l] / NB != bk) {
                            const int item = (int)it_l.size() - lb_ptr[bk];
                            if (!is_to) from_item[l] = item;
                            nl_src[k] = item;
                            it_l.push_back(l); it_a.push_back(from[l]); it_b.push_back(to[l]);
                        }
                    }
                for (int n = n0; n < n1; ++n)
                    for (int k = nl_ptr[n]; k < nl_ptr[n + 1]; ++k) {
                        const int l = nl_ent[k] >> 1;
                        if ((nl_ent[k] & 1) && from[l] / NB == bk) nl_src[k] = from_item[l];
                    }
                max_items = std::max(max_items, it_l.size() - (size_t)lb_ptr[bk]);
            }
            lb_ptr[nblk] = (int)it_l.size();
            link_node_lds = max_items * sizeof(real);
            if (link_node_lds > 65536)
                throw ValueError{"a block of 256 drainage nodes has more links than its LDS holds"};
            {
                const size_t ni = it_l.size();
                std::vector<int> it_a0(ni), it_a1(ni), it_b0(ni), it_b1(ni);
                for (size_t q = 0; q < ni; ++q) {
                    it_a0[q] = ni_ptr[it_a[q]]; it_a1[q] = ni_ptr[it_a[q] + 1];
                    it_b0[q] = ni_ptr[it_b[q]]; it_b1[q] = ni_ptr[it_b[q] + 1];
                }
                const int KB = fh::kRowEntries;
                std::vector<int> nls((size_t)KB * NT, 0);
                for (int n = 0; n < NT; ++n)
                    for (int u = 0; u < KB && nl_ptr[n] + u < nl_ptr[n + 1]; ++u)
                        nls[(size_t)u * NT + n] = nl_src[nl_ptr[n] + u];
                d_lb_ptr.set(lb_ptr, s); d_it_l.set(it_l, s); d_it_a.set(it_a, s); d_it_b.set(it_b, s);
                d_it_a0.set(it_a0, s); d_it_a1.set(it_a1, s); d_it_b0.set(it_b0, s); d_it_b1.set(it_b1, s);
                d_nls.set(nls, s);
            }
            d_nl_src.set(nl_src, s);
            // node -> links: 'from' entries then 'to' entries, each ascending, with the far node
            std::vector<int> nl2_ptr(NT + 1, 0), nl2_ent(nl_ent.size()), nl2_oth(nl_ent.size());
            int w = 0;
            for (int n = 0; n < NT; ++n) {
                nl2_ptr[n] = w;
                for (int pass = 0; pass < 2; ++pass)
                    for (int k = nl_ptr[n]; k < nl_ptr[n + 1]; ++k) {
                        const int en = nl_ent[k];
                        if ((en & 1) != pass) continue;
                        const int l = en >> 1;
                        nl2_ent[w] = en;
                        nl2_oth[w] = pass ? from[l] : to[l];
                        ++w;
                    }
            }
            nl2_ptr[NT] = w;
            std::vector<real> ne_sw(ne_idx.size());
            for (size_t k = 0; k < ne_idx.size(); ++k) ne_sw[k] = sw[ne_idx[k]];
            std::vector<int> rc_enode(rc_eidx.size());
            std::vector<real> rc_esw(rc_eidx.size());
            for (size_t k = 0; k < rc_eidx.size(); ++k) {
                rc_enode[k] = snode[rc_eidx[k]];
                rc_esw[k] = sw[rc_eidx[k]];
            }
            // row-aligned copies of each row's first entries
            {
                const int KB = fh::kRowEntries;
                std::vector<int> nle((size_t)KB * NT, 0), nl2e((size_t)KB * NT, 0), nl2o((size_t)KB * NT, 0);
                std::vector<real> nesw0(NT, 0);
                for (int n = 0; n < NT; ++n) {
                    for (int u = 0; u < KB; ++u) {
                        if (nl_ptr[n] + u < nl_ptr[n + 1]) nle[(size_t)u * NT + n] = nl_ent[nl_ptr[n] + u];
                        if (nl2_ptr[n] + u < nl2_ptr[n + 1]) {
                            nl2e[(size_t)u * NT + n] = nl2_ent[nl2_ptr[n] + u];
                            nl2o[(size_t)u * NT + n] = nl2_oth[nl2_ptr[n] + u];
                        }
                    }
                    if (ne_ptr[n] < ne_ptr[n + 1]) nesw0[n] = ne_sw[ne_ptr[n]];
                }
                const int CI = icl.total;
                if (icl.tile != fh::kInletRows) throw ValueError{"inlet-cell layout tile differs from the inlet kernel's"};
                // per inlet entry: cell, row, node head data, and the node's inlets
                {
                    std::vector<int> ie_cell((size_t)nic, 0), ie_row((size_t)nic, 0), ie_s0((size_t)nic, 0),
                        ie_ns((size_t)nic, 0), ie_sc((size_t)KB * std::max(nic, 1), 0);
                    std::vector<real> ie_inv((size_t)nic, 0), ie_As((size_t)nic, 1), ie_sq((size_t)KB * std::max(nic, 1), 0);
                    size_t max_ent = 1;
                    for (int j = 0; j < CI; ++j) {
                        if (j % fh::kInletRows == 0) {
                            const int jb = std::min(CI, j + fh::kInletRows);
                            max_ent = std::max(max_ent, (size_t)(ic_ptr[jb] - ic_ptr[j]));
                        }
                        for (int k = ic_ptr[j]; k < ic_ptr[j + 1]; ++k) {
                            const int n = ic_node[k];
                            ie_cell[k] = ic_cells[j];
                            ie_row[k] = j;
                            ie_inv[k] = inv[n];
                            ie_As[k] = As[n];
                            ie_s0[k] = ni_ptr[n];
                            ie_ns[k] = ni_ptr[n + 1] - ni_ptr[n];
                            for (int u = 0; u < KB; ++u) {
                                const bool has = u < ie_ns[k];
                                ie_sc[(size_t)u * nic + k] = has ? ni_cell[ni_ptr[n] + u] : ic_cells[j];
                                ie_sq[(size_t)u * nic + k] = has ? ni_surq[ni_ptr[n] + u] : real(0);
                            }
                        }
                    }
                    inlet_lds = (2 * max_ent + fh::kInletRows) * sizeof(real);
                    if (inlet_lds > 65536)
                        throw ValueError{"a block of 256 inlet cells has more inlets than its LDS holds"};
                    d_ie_cell.set(ie_cell, s); d_ie_row.set(ie_row, s); d_ie_s0.set(ie_s0, s); d_ie_ns.set(ie_ns, s);
                    d_ie_sc.set(ie_sc, s); d_ie_inv.set(ie_inv, s); d_ie_As.set(ie_As, s); d_ie_sq.set(ie_sq, s);
                    d_ic_h.set(std::vector<real>((size_t)CI, 0), s);
                }
                std::vector<typename fh::V4<real>::type> lk_a((size_t)LT), lk_b((size_t)LT);
                for (int l = 0; l < LT; ++l) {
                    lk_a[l].x = linv[l]; lk_a[l].y = D[l]; lk_a[l].z = llen[l]; lk_a[l].w = man[l];
                    lk_b[l].x = minor[l]; lk_b[l].y = lcap[l]; lk_b[l].z = rw[l]; lk_b[l].w = rh[l];
                }
                const int CR = rcl.total;
                std::vector<int> rcf_node(CR, 0), rcf_fb(CR, 0);
                std::vector<real> rcf_sw(CR, 0);
                for (int j = 0; j < CR; ++j) {
                    if (rc_eptr[j] < rc_eptr[j + 1]) { rcf_node[j] = rc_enode[rc_eptr[j]]; rcf_sw[j] = rc_esw[rc_eptr[j]]; }
                    if (rc_fptr[j] < rc_fptr[j + 1]) rcf_fb[j] = rc_fidx[rc_fptr[j]];
                }
                d_nle.set(nle, s); d_nl2e.set(nl2e, s);
                d_nl2o.set(nl2o, s); d_nesw0.set(nesw0, s);
                d_lk_a.set(lk_a, s); d_lk_b.set(lk_b, s);
                d_rcf_node.set(rcf_node, s); d_rcf_sw.set(rcf_sw, s); d_rcf_fb.set(rcf_fb, s);
                std::vector<real> ic_area(CI, 1), rc_area(CR, 1);
                for (int j = 0; j < CI; ++j) ic_area[j] = area_of_cell[(size_t)ic_cells[j]];
                for (int j = 0; j < CR; ++j) rc_area[j] = area_of_cell[(size_t)rc_cells[j]];
                d_ic_area.set(ic_area, s); d_rc_area.set(rc_area, s);
                // return rows with one fallback node and nothing else go to that node's
                // thread in the node update; the others stay with the return kernel
                std::vector<int> nr_cell(NT, -1);
                std::vector<real> nr_area(NT, 1);
                std::vector<std::vector<int>> rx(M);
                for (int i = 0; i < M; ++i)
                    for (int j = rcl.base[i]; j < rcl.base[i] + rcl.count[i]; ++j) {
                        const bool simple = rc_eptr[j] == rc_eptr[j + 1] && rc_fptr[j + 1] - rc_fptr[j] == 1 &&
                                            rc_rptr[j] == rc_rptr[j + 1];
                        if (simple) {
                            const int n = rc_fidx[rc_fptr[j]];
                            nr_cell[n] = rc_cells[j];
                            nr_area[n] = rc_area[j];
                        } else {
                            rx[i].push_back(j);
                        }
                    }
                rxl.build(counts_of(rx), 256);
                std::vector<int> rx_rows((size_t)rxl.total, 0);
                for (int i = 0; i < M; ++i)
                    for (size_t k = 0; k < rx[i].size(); ++k) rx_rows[(size_t)rxl.base[i] + k] = rx[i][k];
                d_rx_rows.set(rx_rows, s); d_nr_cell.set(nr_cell, s); d_nr_area.set(nr_area, s);
                seg_rx.set(rxl, order, s);
            }
            d_ni_cell.set(ni_cell, s); d_ni_surq.set(ni_surq, s);
            std::vector<real> zS((size_t)nslots, 0);
            d_ni_net.set(zS, s);
            n_inlet_slots = nslots;
            d_ic_node.set(ic_node, s); d_ic_slot.set(ic_slot, s); d_ic_capq.set(ic_capq, s);
            d_nl2_ptr.set(nl2_ptr, s); d_nl2_ent.set(nl2_ent, s); d_nl2_oth.set(nl2_oth, s);
            d_ne_sw.set(ne_sw, s);
            d_rc_enode.set(rc_enode, s); d_rc_esw.set(rc_esw, s);
        }
        d_ni_ptr.set(ni_ptr, s); d_nl_ptr.set(nl_ptr, s); d_nl_ent.set(nl_ent, s);
        d_ne_ptr.set(ne_ptr, s);
        d_ic_cells.set(ic_cells, s); d_ic_ptr.set(ic_ptr, s);
        d_geom.set(geom, s); d_Q.set(Q, s);
        std::vector<real> zL(LT, 0);
        d_req.set(zL, s);
        d_snorm.set(by_id(snorm, order), s);
        d_rc_cells.set(rc_cells, s); d_rc_eptr.set(rc_eptr, s);
        d_rc_fptr.set(rc_fptr, s); d_rc_fidx.set(rc_fidx, s); d_rc_rptr.set(rc_rptr, s); d_rc_ridx.set(rc_ridx, s);
        d_rS.set(rS, s); d_rcc.set(rcc, s); d_rrc.set(rrc, s); d_rstop.set(rstop, s); d_rstore.set(rstore, s);
        d_ext_q.set(q_all, s); d_rcv_qbase.set(rcv_qbase, s); d_rcv_stride.set(rcv_stride, s);
        d_ext_ov.set(std::vector<float>((size_t)std::max(Xt, 1), 0.0f), s);
        d_xc_cells.set(xc_cells, s); d_xc_ptr.set(xc_ptr, s); d_xc_idx.set(xc_idx, s);
        seg_nodes.set(nodes, order, s); seg_links.set(links, order, s); seg_icl.set(icl, order, s);
        seg_rcl.set(rcl, order, s); seg_xcl.set(xcl, order, s);
        d_part_inlet.set(std::vector<double>(2 * (size_t)std::max(icl.blocks, 1), 0.0), s);
        d_part_node.set(std::vector<double>(3 * (size_t)std::max(nodes.blocks, 1), 0.0), s);
        d_part_ret.set(std::vector<double>(2 * (size_t)std::max(rxl.blocks, 1), 0.0), s);
        d_part_ext.set(std::vector<double>((size_t)std::max(xcl.blocks, 1), 0.0), s);
        n_receivers = Xt;
        // device view
        fh::DrainDev<real>& d = dd;
        std::memset(&d, 0, sizeof(d));
        d.N = NT; d.I = It; d.Ci = icl.total; d.L = LT; d.S = St; d.Cr = rcl.total; d.R = Rt; d.Rcv = Xt;
        d.Cx = xcl.total;
        d.seg_node = seg_nodes.seg(); d.seg_link = seg_links.seg(); d.seg_icell = seg_icl.seg();
        d.seg_ret = seg_rcl.seg(); d.seg_ext = seg_xcl.seg();
        d.seg_rx = seg_rx.seg(); d.rx_rows = d_rx_rows.p; d.nr_cell = d_nr_cell.p; d.nr_area = d_nr_area.p;
        d.inv = d_inv.p; d.As = d_As.p; d.dmax = d_dmax.p; d.nflag = d_nflag.p; d.outfall = d_outfall.p;
        d.V = d_V.p; d.Vp = d_Vp.p; d.sur_factor = d_sfac.p;
        d.overflow = d_over.p;
        d.fallback = d_fb.p; d.wsum = d_wsum.p; d.store_returns = rxl.blocks > 0 ? 1 : 0;
        d.ni_ptr = d_ni_ptr.p; d.ni_cell = d_ni_cell.p; d.ni_surq = d_ni_surq.p;
        d.ni_net = d_ni_net.p; d.nl_ptr = d_nl_ptr.p; d.nl_ent = d_nl_ent.p;
        d.nl2_ptr = d_nl2_ptr.p; d.nl2_ent = d_nl2_ent.p; d.nl2_oth = d_nl2_oth.p;
        d.ne_ptr = d_ne_ptr.p; d.ne_sw = d_ne_sw.p;
        d.ic_cells = d_ic_cells.p; d.ic_ptr = d_ic_ptr.p; d.ic_node = d_ic_node.p; d.ic_slot = d_ic_slot.p;
        d.ic_capq = d_ic_capq.p;
        d.nle = d_nle.p; d.nl2e = d_nl2e.p; d.nl2o = d_nl2o.p;
        d.lb_ptr = d_lb_ptr.p; d.it_l = d_it_l.p; d.it_a = d_it_a.p; d.it_b = d_it_b.p;
        d.it_a0 = d_it_a0.p; d.it_a1 = d_it_a1.p; d.it_b0 = d_it_b0.p; d.it_b1 = d_it_b1.p;
        d.nl_src = d_nl_src.p; d.nls = d_nls.p;
        d.ie_cell = d_ie_cell.p; d.ie_row = d_ie_row.p; d.ie_s0 = d_ie_s0.p; d.ie_ns = d_ie_ns.p;
        d.ie_sc = d_ie_sc.p; d.ie_inv = d_ie_inv.p; d.ie_As = d_ie_As.p; d.ie_sq = d_ie_sq.p;
        d.ic_h = d_ic_h.p; d.lk_a = d_lk_a.p; d.lk_b = d_lk_b.p;
        d.nesw0 = d_nesw0.p;
        d.rcf_node = d_rcf_node.p; d.rcf_sw = d_rcf_sw.p; d.rcf_fb = d_rcf_fb.p;
        d.ic_area = d_ic_area.p; d.rc_area = d_rc_area.p;
        d.geom = d_geom.p; d.Q = d_Q.p; d.req = d_req.p; d.snorm = d_snorm.p;
        d.rc_cells = d_rc_cells.p; d.rc_eptr = d_rc_eptr.p; d.rc_enode = d_rc_enode.p; d.rc_esw = d_rc_esw.p;
        d.rc_fptr = d_rc_fptr.p;
        d.rc_fidx = d_rc_fidx.p; d.rc_rptr = d_rc_rptr.p; d.rc_ridx = d_rc_ridx.p;
        d.rS = d_rS.p; d.rcc = d_rcc.p; d.rrc = d_rrc.p; d.rstop = d_rstop.p; d.rstore = d_rstore.p;
        d.ext_q = d_ext_q.p; d.rcv_qbase = d_rcv_qbase.p; d.rcv_stride = d_rcv_stride.p;
        d.ext_q_override = d_ext_ov.p; d.use_override = 0;
        d.xc_cells = d_xc_cells.p; d.xc_ptr = d_xc_ptr.p; d.xc_idx = d_xc_idx.p;
    }

    // ------------------------------------------------------------------ stepping
    fh::ForcingTables forcing_tables() const {
        return fh::ForcingTables{d_rain_off.p, d_rain_n.p, d_ext_off.p, d_ext_n.p,
                                 d_rain_s.p, d_rain_e.p, d_rain_r.p, d_ext_t.p};
    }

    // Stage launches go through here.  While a kernel is being timed the launch
    // carries start/stop events in its dispatch packet (hipExtLaunchKernel), so
    // the measured interval is the kernel's own execution, as a profiler sees it.
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    template <typename... KA, typename... A>
    void go(void (*kernel)(KA...), unsigned grid, unsigned block, A... args) {
        go_lds(kernel, grid, block, 0, args...);
    }
    template <typename... KA, typename... A>
    void go_lds(void (*kernel)(KA...), unsigned grid, unsigned block, size_t lds, A... args) {
        if (ev_start)
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, ev_start, ev_stop, 0u,
                                  static_cast<KA>(args)...);
        else
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, static_cast<KA>(args)...);
    }

    // CFL minimum over every cell of each member into cfl_min (stable_timestep)
    void launch_cfl() {
        if (li)
            go(fh::k_cfl<real, true>, begin_l.blocks, 256, kp, par, d_state.p, d_lc.p,
               fh::CellRanges{seg_begin.seg(), d_cell0.p, d_cell1.p});
        else
            go(fh::k_cfl<real, false>, begin_l.blocks, 256, kp, par, d_state.p, d_lc.p,
               fh::CellRanges{seg_begin.seg(), d_cell0.p, d_cell1.p});
        go(fh::k_cfl_final, (unsigned)M, 64, ctl, kp.cfl_slots, par);
    }
    // local-inertial surface: the step's CFL pass, edge discharges, donor limiting
    void launch_li_cfl() {
        go(fh::k_li_cfl<real>, begin_l.blocks, 256, ctl, kp, par, d_state.p, d_lc.p,
           fh::CellRanges{seg_begin.seg(), d_cell0.p, d_cell1.p});
    }
    void launch_li_flux() {
        if (edges.blocks == 0) return;
        fh::EdgeSeg es{edge_seg(), edges.blocks, seg_orph.seg(), d_ocells.p};
#define FH_LIFLUX(SH_, G_)                                                                                 \
        go(fh::k_li_flux<real, SH_, G_>, edges.blocks, bs, ctl, kp, es, par, forcing_tables(), d_el.p, d_er.p,   \
           d_nx.p, d_ny.p, d_len.p, d_state.p, d_lc.p, d_cell_n.p, reinterpret_cast<fh::Rec3<real>*>(d_rec.p), \
           d_qli.p)
        if (shared) FH_LIFLUX(true, true); else if (M > 1) FH_LIFLUX(false, true); else FH_LIFLUX(false, false);
#undef FH_LIFLUX
    }
    void launch_li_donor() {
#define FH_DONOR(SH_, G_)                                                                                  \
        go(fh::k_li_donor<real, SH_, G_>, cells.blocks, bs, ctl, kp, cell_seg(), par, forcing_tables(), d_ce0.p, \
           d_ce1.p, d_ce2.p, d_perm.p, d_qli.p, d_recR.p, d_state.p, d_area.p)
        if (shared) FH_DONOR(true, true); else if (M > 1) FH_DONOR(false, true); else FH_DONOR(false, false);
#undef FH_DONOR
    }
    void launch_flux() {
        if (edges.blocks + orphans.blocks == 0) return;
        fh::EdgeSeg es{edge_seg(), edges.blocks, seg_orph.seg(), d_ocells.p};
        const unsigned grid = (unsigned)(edges.blocks + orphans.blocks);
#define FH_FLUX(EPT_, SH_, G_)                                                                                \
        go(fh::k_flux<real, EPT_, SH_, G_>, grid, bs, ctl, kp, es, par, d_el.p, d_er.p, d_nx.p, d_ny.p, d_len.p, \
           d_state.p, d_lc.p, d_rec.p, d_recR.p, d_stage_h.p)
#define FH_FLUX_S(EPT_) \
        if (shared) FH_FLUX(EPT_, true, true); else if (M > 1) FH_FLUX(EPT_, false, true); else FH_FLUX(EPT_, false, false)
        if (ept == 4) { FH_FLUX_S(4); }
        else if (ept == 2) { FH_FLUX_S(2); }
        else { FH_FLUX_S(1); }
#undef FH_FLUX_S
#undef FH_FLUX
    }
    void launch_update() {
        const fh::CellSources<real> src{d_n2.p, d_rainm.p, d_f0.p, d_fc.p, d_k.p, d_cls.p, d_cls_tab.p, d_cls_hk.p};
        const fh::Seg cs = cell_seg();
#define FH_UPDATE(COMP_, CLS_, LI_, SH_, G_)                                                                \
        go(fh::k_update<real, COMP_, CLS_, LI_, SH_, G_>, cells.blocks, bs, ctl, kp, cs, par,                        \
           forcing_tables(), d_ce0.p, d_ce1.p, d_ce2.p,                                                         \
           d_rec.p, d_recR.p, d_state.p, d_comp.p, d_wet.p, d_area.p, d_lc.p, src, d_part_upd.p)
#define FH_UPDATE_S(COMP_, CLS_, LI_) \
        if (shared) { FH_UPDATE(COMP_, CLS_, LI_, true, true); }                                             \
        else if (M > 1) { FH_UPDATE(COMP_, CLS_, LI_, false, true); }                                        \
        else { FH_UPDATE(COMP_, CLS_, LI_, false, false); }
#define FH_UPDATE_M(COMP_, CLS_) \
        if (li) { FH_UPDATE_S(COMP_, CLS_, true); } else { FH_UPDATE_S(COMP_, CLS_, false); }
        if (P.compensated) {
            if (use_classes) { FH_UPDATE_M(true, true); } else { FH_UPDATE_M(true, false); }
        } else {
            if (use_classes) { FH_UPDATE_M(false, true); } else { FH_UPDATE_M(false, false); }
        }
#undef FH_UPDATE_M
#undef FH_UPDATE_S
#undef FH_UPDATE
    }

    template <class Timed>
    void launch_drainage(Timed&& timed) {
        const hipStream_t s = stream;
        const double rb = sizeof(real);
        const fh::DrainDev<real> d = dd;
        if (xcl.blocks > 0)
            timed("external", (double)d.Rcv * 12.0 + d.Cx * 16.0, [&] {
                go(fh::k_external<real>, xcl.blocks, 256, ctl, kp, d, d_state.p, d_area.p, d_part_ext.p);
            });
        if (nodes.blocks == 0) return;
        if (icl.blocks > 0)
            timed("inlet", (double)d.Ci * (12 + 4 * rb) + (double)d.I * (8 + 5 * rb), [&] {
                if (M > 1)
                    go_lds(fh::k_inlet<real, true>, icl.blocks, fh::kInletThreads, inlet_lds, ctl, kp, d, d_state.p,
                           d_part_inlet.p);
                else
                    go_lds(fh::k_inlet<real, false>, icl.blocks, fh::kInletThreads, inlet_lds, ctl, kp, d, d_state.p,
                           d_part_inlet.p);
            });
        timed("link_node", (double)d.N * (8 + 4 * rb) + (double)d.L * (8 + 11 * rb), [&] {
            const unsigned grid = nodes.blocks + (icl.total + fh::kLinkThreads - 1) / fh::kLinkThreads;
            if (M > 1)
                go_lds(fh::k_link_node<real, true>, grid, fh::kLinkThreads, link_node_lds, ctl, kp, d, nodes.blocks,
                       d_state.p);
            else
                go_lds(fh::k_link_node<real, false>, grid, fh::kLinkThreads, link_node_lds, ctl, kp, d, nodes.blocks,
                       d_state.p);
        });
        timed("node_update", (double)d.N * (16 + 14 * rb) + (double)d.L * (16 + 3 * rb), [&] {
            if (M > 1) go(fh::k_node_update<real, true>, nodes.blocks, fh::kNodeBlock, ctl, kp, d, d_state.p, d_part_node.p);
            else go(fh::k_node_update<real, false>, nodes.blocks, fh::kNodeBlock, ctl, kp, d, d_state.p, d_part_node.p);
        });
        if (rxl.blocks > 0)
            timed("return", (double)d.Cr * (16 + 3 * rb) + (double)d.N * 2 * rb, [&] {
                if (M > 1) go(fh::k_return<real, true>, rxl.blocks, 256, ctl, kp, d, d_state.p, d_area.p, d_part_ret.p);
                else go(fh::k_return<real, false>, rxl.blocks, 256, ctl, kp, d, d_state.p, d_area.p, d_part_ret.p);
            });
    }

    // One step: flux and update on the clock of parity `par`; the update writes
    // the next step's clock into the other slot, so the parity flips per step.
    // cfl_pass: the local-inertial CFL pass is launched (a single-scenario graph
    // captured for its step phase leaves it out of the steps that cannot evaluate)
    void launch_step(bool cfl_pass = true) {
        if (li) {
            if (cfl_pass) launch_li_cfl();
            launch_li_flux();
            launch_li_donor();
        } else {
            launch_flux();
        }
        launch_update();
        launch_drainage([](const char*, double, auto&& fn) { fn(); });
        par ^= 1;
        ctl_fresh = false;
        wet_pending = true;
    }

    // Captured step graphs of the full-step path, one per (steps, parity of the
    // first step), for the chunk size K and every power of two below it: an
    // advance of n steps is n / K chunks of K plus the binary decomposition of the
    // rest, so a step budget below K takes at most log2(K) + 1 launches and no
    // no-op steps.  All are captured at the first advance (capture records the
    // launches and runs nothing).
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int steps = 0;
    };
    std::map<int, StepGraph> graphs;   // key (steps << 1) | parity
    // single-scenario local-inertial graphs, keyed by (full / short, parity, step
    // phase mod the CFL interval): only the steps that can evaluate launch the CFL pass
    std::map<int64_t, StepGraph> li_graphs;
    int64_t host_step = 0;    // the member's step count at the next chunk (phase-keyed graphs)
    bool ctl_fresh = false;   // ctl_h holds the device control blocks (no launch since the last copy)
    bool wet_pending = false; // the last executed step's wet-time increment is not applied yet

    bool li_phased() const { return li && M == 1 && P.use_graph; }

    static void destroy(StepGraph& g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g = StepGraph{};
    }

    void drop_graph() {
        for (auto& kv : graphs) destroy(kv.second);
        graphs.clear();
        for (auto& kv : li_graphs) destroy(kv.second);
        li_graphs.clear();
    }

    // Chunk sizes of the full-step graphs: K, then the powers of two below K.
    std::vector<int> chunk_sizes() const {
        const int K = std::max(1, P.steps_per_launch);
        std::vector<int> v{K};
        int q = 1;
        while (q * 2 < K) q *= 2;
        for (; q >= 1; q /= 2)
            if (q < K) v.push_back(q);
        return v;
    }

    void prebuild_graphs() {
        if (!P.use_graph || li_phased() || !graphs.empty()) return;
        for (int k : chunk_sizes())
            for (int q = 0; q < 2; ++q) {
                const int p0 = par;
                par = q;
                build_graph(graphs[(k << 1) | q], k);
                par = p0;
            }
    }

    // Launches n steps: chunks of K, then the largest chunk sizes that fit.
    void launch_steps(int64_t n) {
        const std::vector<int> sizes = chunk_sizes();
        if (li_phased()) {
            const int K = sizes[0], Kt = std::min(K, kTailSteps);
            while (n >= K) { launch_chunk(K); n -= K; }
            while (n > 0) { launch_chunk(Kt); n -= Kt; }
            return;
        }
        for (int k : sizes)
            while (n >= k) { launch_chunk(k); n -= k; }
    }

    // phase < 0: every step launches the CFL pass
    void build_graph(StepGraph& g, int k, int64_t phase = -1) {
        destroy(g);
        const int p0 = par;
        FH_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < k; ++i)
            launch_step(phase < 0 || (kp.fixed_dt <= 0.0 && (phase + i) % kp.interval == 0));
        FH_CHECK(hipStreamEndCapture(stream, &g.graph));
        par = p0;   // capture only recorded the steps
        FH_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        // move the graph's device-side setup out of its first launch
        FH_CHECK(hipGraphUpload(g.exec, stream));
        g.steps = k;
    }

    void launch_chunk(int k) {
        if (li_phased()) {
            // steps past the target are no-ops and do not count, so host_step is exact
            // for every step that runs
            const int64_t phase = host_step % kp.interval;
            const int64_t key = ((int64_t)(k == std::max(1, P.steps_per_launch) ? 0 : 1) * 2 + par) *
                                    (int64_t)kp.interval + phase;
            StepGraph& g = li_graphs[key];
            if (!g.exec || g.steps != k) build_graph(g, k, phase);
            FH_CHECK(hipGraphLaunch(g.exec, stream));
            par ^= (k & 1);
            host_step += k;
        } else if (P.use_graph) {
            StepGraph& g = graphs[(k << 1) | par];
            if (!g.exec || g.steps != k) build_graph(g, k);
            FH_CHECK(hipGraphLaunch(g.exec, stream));
            par ^= (k & 1);
        } else {
            for (int i = 0; i < k; ++i) launch_step();
        }
        wet_pending = true;
    }

    // Sum every member's ledger accumulation slots into its FP64 ledger; with
    // check, also flag non-finite depths and node volumes (one launch).
    void fold(bool check = false) {
        fh::FoldTables f{};
        auto group = [&](int i, double* acc, const DevSeg& seg, int n, int t0, int t1, int t2) {
            f.g[i] = fh::FoldGroup{acc, seg.mem_blk0.p, seg.mem_nblk.p, n, {t0, t1, t2, 0}};
        };
        f.g[0] = fh::FoldGroup{d_part_upd.p, seg_cells.mem_blk0.p, seg_cells.mem_nblk.p, 4,
                               {fh::L_RAIN, fh::L_INFIL, fh::L_BIN, fh::L_BOUT}};
        group(1, d_part_inlet.p, seg_icl, 2, fh::L_CAPTURE, fh::L_SURCHARGE, 0);
        group(2, d_part_node.p, seg_nodes, 3, fh::L_OUTFALL, fh::L_OVERFLOW, fh::L_UNRETURNED);
        group(3, d_part_ret.p, seg_rx, 2, fh::L_RCAPTURE, fh::L_RLOSS, 0);
        group(4, d_part_ext.p, seg_xcl, 1, fh::L_EXTERNAL, 0, 0);
        f.ngroups = 5;
        launch_fold(f, check);
    }

    // the fold blocks, plus (check) blocks that flag a non-finite depth or node
    // volume of the batch in ctls[0].nonfinite (cleared by k_set_target)
    void launch_fold(const fh::FoldTables& f, bool check) {
        const int64_t items = check ? (int64_t)cells.total + nodes.total : 0;
        const int extra = (int)((items + fh::kFoldThreads - 1) / fh::kFoldThreads);
        fh::k_fold<real><<<M * f.ngroups + extra, fh::kFoldThreads, 0, stream>>>(
            ctl, f, M, check ? cells.total : 0, d_state.p, check ? nodes.total : 0, d_V.p);
        ctl_fresh = false;
    }

    void finalize_pending() {
        fh::k_finalize_wet<real><<<std::max(cells.blocks, 1), bs, 0, stream>>>(ctl, kp, seg_cells.seg(), par,
                                                                               d_state.p, d_wet.p);
        fh::k_clear_pending<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, par);
        ctl_fresh = false;
    }

    // The wet-time increment of the last executed step is applied by the next
    // step's update, or here, before the host reads or replaces the state.
    void finalize_if_pending() {
        if (!wet_pending || !packed || dirty) return;
        finalize_pending();
        wet_pending = false;
    }

    bool any_running(double target, int64_t max_steps) const {
        for (int m = 0; m < M; ++m)
            if (ctl_h[m].clk[par].t < target && ctl_h[m].clk[par].interval_step < max_steps) return true;
        return false;
    }

    int advance_to(double target, double final_t, int64_t max_steps) override {
        ensure_packed();
        if (max_steps <= 0) throw ValueError{"max_steps must be positive"};
        prebuild_graphs();
        fh::k_set_target<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, par, target, final_t, max_steps);
        const int K = std::max(1, P.steps_per_launch);
        if (!ctl_fresh) pull_ctl();
        // the host's copy of the clocks is current; the step counters restart
        for (int m = 0; m < M; ++m) ctl_h[m].clk[par].interval_step = 0;
        bool done = false;
        while (!done) {
            host_step = ctl_h[0].clk[par].step;
            // steps the running members still need: their step budget, or the time
            // left at the last dt (the dt cap before any step, which never
            // overshoots).  A batch that covers every member's need is expected to
            // end the advance, so the ledger fold and the finiteness check are
            // queued ahead of its readback; a larger need is stepped in batches
            // of three quarters of it, since the dt can grow at the next CFL
            // evaluation and steps past the target are no-ops.
            double need = 0.0;
            bool exact = true;
            for (int m = 0; m < M; ++m) {
                const fh::Clock& ck = ctl_h[m].clk[par];
                if (!(ck.t < target && ck.interval_step < max_steps)) continue;
                const double left = (double)(max_steps - ck.interval_step);
                const double dt_last = ctl_h[m].dt > 0 ? ctl_h[m].dt : (kp.fixed_dt > 0.0 ? kp.fixed_dt : P.max_dt);
                const double by_time = std::ceil((target - ck.t) / std::max(dt_last, 1e-9));
                if (by_time < left) exact = false;
                need = std::max(need, std::min(left, by_time));
            }
            if (need <= 0.0) break;
            int64_t n;
            bool last;
            if (exact || need <= K) {
                n = (int64_t)std::min(need, 64.0 * K);
                last = need <= 64.0 * K;
            } else {
                n = (int64_t)std::min(16.0 * K, std::max((double)K, std::floor(0.75 * need)));
                last = false;
            }
            launch_steps(n);
            if (last) fold(true);
            pull_ctl();
            done = !any_running(target, max_steps);
            if (done && !last) { fold(true); pull_ctl(); }
        }
        for (int m = 0; m < M; ++m)
            if (ctl_h[m].li_missed) throw std::runtime_error("local-inertial step evaluated the CFL condition "
                                                             "without its CFL pass");
        if (ctl_h[0].nonfinite) return FH_ERR_SIMULATION;
        for (int m = 0; m < M; ++m)
            if (ctl_h[m].clk[par].t < target) return FH_ERR_BUDGET;
        return FH_OK;
    }

    void step_once(double dt, double rain, const float* q, double* diag) override {
        ensure_packed();
        if (!(dt > 0.0)) throw ValueError{"timestep_s must be positive."};
        if (n_receivers > 0) {
            std::vector<float> ov((size_t)n_receivers, 0.0f);
            // the override table holds the selected member's receivers (single-member use)
            if (q) {
                const int xb = 0;
                for (int64_t j = 0; j < cm().Rcv; ++j) ov[xb + j] = q[j];
            }
            d_ext_ov.set(ov, stream);
            dd.use_override = 1;
        }
        fold();   // commits anything pending, so the next fold holds this step alone
        fh::k_set_target<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, par, INFINITY, INFINITY, 1);
        fh::k_set_override<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, 1, dt, rain);
        launch_step();
        fh::k_set_override<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, 0, 0.0, 0.0);
        fold();
        dd.use_override = 0;
        pull_ctl();
        for (int m = 0; m < M; ++m) {
            fh::Clock& ck = ctl_h[m].clk[par];
            ck.eval = (P.fixed_dt <= 0.0 && (ck.step % kp.interval) == 0) ? 1 : 0;
        }
        push_ctl();
        if (diag)
            for (int k = 0; k < fh::kLedgerTerms; ++k) diag[k] = ctl_h[cur].diag[k];
    }

    void dry_weather_step(double dt, const float* inflow, const float* tw, int cap, double* diag) override {
        ensure_packed();
        ctl_fresh = false;
        if (M != 1) throw ValueError{"dry-weather stepping acts on single-scenario contexts"};
        if (!(dt > 0.0)) throw ValueError{"timestep_s must be positive."};
        const Member<real>& m = mem[0];
        const real dtr = (real)dt;
        double inflow_sum = 0.0;
        if (inflow) {
            std::vector<real> v((size_t)std::max(nodes.total, 1), real(0));
            for (int64_t n = 0; n < m.N; ++n) {
                const real q = (real)inflow[n];
                v[(size_t)(nodes.base[0] + m.np((int)n))] = q;
                const real qm = (q != q || q > real(0)) ? q : real(0);
                inflow_sum += (double)(qm * dtr);
            }
            d_dw_in.set(v, stream);
        }
        if (tw) {
            std::vector<real> v((size_t)std::max(nodes.total, 1), real(0));
            for (int64_t n = 0; n < m.N; ++n) {
                if (!std::isfinite(tw[n])) throw ValueError{"outfall tailwater heads must be finite"};
                v[(size_t)(nodes.base[0] + m.np((int)n))] = (real)tw[n];
            }
            d_dw_tw.set(v, stream);
        }
        fold();
        pull_ctl();
        const fh::Ctl saved = ctl_h[0];     // the run's dt, flags and ledger, restored below
        if (nodes.blocks > 0) {
            const int threads = std::max(nodes.total, n_inlet_slots);
            fh::k_dw_prep<real><<<(threads + 255) / 256, 256, 0, stream>>>(ctl, dd, n_inlet_slots,
                                                                          inflow ? d_dw_in.p : nullptr, dt);
            fh::DrainDev<real> d = dd;
            d.dry = 1;
            d.dry_cap = cap ? 1 : 0;
            d.dry_tw = tw ? d_dw_tw.p : nullptr;
            go_lds(fh::k_link_node<real, false>, nodes.blocks, fh::kLinkThreads, link_node_lds, ctl, kp, d, nodes.blocks,
                   d_state.p);
            go(fh::k_node_update<real, false>, nodes.blocks, fh::kNodeBlock, ctl, kp, d, d_state.p, d_part_node.p);
        }
        fold();
        pull_ctl();
        double final_volume = 0.0;
        if (m.N > 0) {
            std::vector<real> V((size_t)nodes.total);
            d_V.get(V.data(), 0, (size_t)nodes.total, stream);
            FH_CHECK(hipStreamSynchronize(stream));
            for (int64_t n = 0; n < m.N; ++n) final_volume += (double)V[(size_t)(nodes.base[0] + m.np((int)n))];
        }
        // the node kernel's slots carry outfall, overflow and the moved volume in dry mode
        diag[0] = inflow_sum;
        diag[1] = ctl_h[0].diag[fh::L_UNRETURNED];
        diag[2] = ctl_h[0].diag[fh::L_OUTFALL];
        diag[3] = ctl_h[0].diag[fh::L_OVERFLOW];
        diag[4] = final_volume;
        std::memcpy(ctl_h[0].ledger, saved.ledger, sizeof(saved.ledger));
        std::memcpy(ctl_h[0].diag, saved.diag, sizeof(saved.diag));
        ctl_h[0].dt = saved.dt;
        ctl_h[0].dt_real = saved.dt_real;
        ctl_h[0].active = saved.active;
        push_ctl();
    }

    double stable_timestep() override {
        ensure_packed();
        launch_cfl();
        pull_ctl();
        return ctl_h[cur].cfl_min;
    }

    std::vector<KernelTiming> profile(int steps) override {
        ensure_packed();
        std::vector<KernelTiming> out;
        // every launch carries start/stop events in its dispatch packet; the launches
        // run back to back as in a step, and the events are read after the last one.
        // The events take no system-scope fence (the stream synchronization before
        // they are read orders them), so, as between the kernels of a step graph, the
        // L2 is neither written back nor invalidated between the timed launches.
        struct Pending { std::string name; double bytes; hipEvent_t a, b; };
        std::vector<Pending> pending;
        auto timed = [&](const char* name, double bytes, auto&& fn) {
            hipEvent_t a, b;
            FH_CHECK(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
            FH_CHECK(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
            ev_start = a;
            ev_stop = b;
            fn();
            ev_start = ev_stop = nullptr;
            FH_CHECK(hipGetLastError());
            pending.push_back(Pending{name, bytes, a, b});
        };
        double Ct = 0, Et = 0;
        for (const auto& m : mem) { Ct += (double)m.C; Et += (double)m.E; }
        const double R4 = sizeof(real), I4 = 4.0;
        // bytes each kernel must move (each required element once per step)
        // (local-inertial: the edge records are {Q, Q nx, Q ny} plus Q and T apart,
        // and the discharge gathers Manning n and the CFL length of both cells)
        const double flux_bytes = Et * (2 * I4 + 3 * R4) + Ct * 4 * R4 + Et * 5 * R4;
        const double li_flux_bytes = Et * (2 * I4 + 3 * R4) + Ct * 6 * R4 + Et * 4 * R4;
        const double upd_bytes = Et * (li ? 4 : 5) * R4 + Ct * (3 * I4 + 8 * R4 + 2 * R4 + 4 * R4 + 2 * R4) +
                                 (P.compensated ? Ct * 8 * R4 : 0.0);
        fh::k_set_target<<<(M + 255) / 256, 256, 0, stream>>>(ctl, M, par, INFINITY, INFINITY, (int64_t)steps + 8);
        for (int i = 0; i < steps; ++i) {
            if (li) {
                timed("li_cfl", 0.0, [&] { launch_li_cfl(); });
                timed("li_flux", li_flux_bytes, [&] { launch_li_flux(); });
                timed("li_donor", Ct * (3 * I4 + 1 + 2 * R4) + Et * 2 * R4, [&] { launch_li_donor(); });
            } else {
                timed("flux", flux_bytes, [&] { launch_flux(); });
            }
            timed("update", upd_bytes, [&] { launch_update(); });
            launch_drainage(timed);
            par ^= 1;
        }
        fold();
        finalize_pending();
        wet_pending = false;
        ctl_fresh = false;
        FH_CHECK(hipStreamSynchronize(stream));
        for (const Pending& q : pending) {
            float ms = 0.f;
            FH_CHECK(hipEventElapsedTime(&ms, q.a, q.b));
            (void)hipEventDestroy(q.a);
            (void)hipEventDestroy(q.b);
            auto it = std::find_if(out.begin(), out.end(), [&](const KernelTiming& k) { return k.name == q.name; });
            if (it == out.end()) {
                out.push_back(KernelTiming{q.name, 0.0, q.bytes, 0});
                it = out.end() - 1;
            }
            it->total_ms += ms;
            it->launches += 1;
        }
        return out;
    }

    int64_t device_bytes() const override {
        int64_t b = 0;
        b += d_el.bytes() + d_er.bytes() + d_nx.bytes() + d_ny.bytes() + d_len.bytes();
        b += d_rec.bytes() + d_recR.bytes() + d_state.bytes() + d_wet.bytes() + d_comp.bytes();
        b += d_area.bytes() + d_n2.bytes() + d_lc.bytes() + d_rainm.bytes() + d_f0.bytes() + d_fc.bytes() + d_k.bytes();
        b += d_cls.bytes() + d_cls_tab.bytes() + d_cls_hk.bytes();
        b += d_ce0.bytes() + d_ce1.bytes() + d_ce2.bytes();
        // drainage: the node, slot, entry and inlet tables by their representative
        // arrays, the links by their two records, flow, request, code and work items
        b += d_inv.bytes() * 12 + d_ni_cell.bytes() * 4 + d_nl_ent.bytes() * 3 + d_ic_node.bytes() * 3 +
             d_lk_a.bytes() + d_lk_b.bytes() + d_Q.bytes() * 2 + d_geom.bytes() + d_it_l.bytes() * 7;
        return b;
    }

    // ------------------------------------------------------------------ members
    fh_params P;
    int M = 1;
    int cur = 0;
    std::vector<Member<real>> mem;
    std::vector<int> order, member_pos, rrow_base;
    fh::KP<real> kp{};
    hipStream_t stream = nullptr;
    fh::Ctl* ctl = nullptr;     // device, [M] indexed by member id
    fh::Ctl* ctl_h = nullptr;   // pinned host mirror
    int bs = 256, ept = 1;
    int par = 0;                // parity of the clock slot the next step runs on
    bool packed = false, dirty = true, any_drainage = false;
    int n_receivers = 0;
    Layout cells, edges, nodes, links, icl, rcl, xcl, begin_l, orphans, rxl;
    DevSeg seg_cells, seg_edges, seg_begin, seg_nodes, seg_links, seg_icl, seg_rcl, seg_xcl, seg_orph, seg_rx;
    DevArray<int> d_rx_rows, d_nr_cell;
    DevArray<real> d_nr_area;
    DevArray<int> d_cell0, d_cell1, d_rain_off, d_rain_n, d_ext_off, d_ext_n, d_ocells;
    DevArray<int> d_el, d_er, d_ce0, d_ce1, d_ce2;
    DevArray<real> d_cell_n;          // local-inertial: Manning n per cell
    DevArray<unsigned char> d_perm;   // local-inertial: entry slots in ascending edge order
    DevArray<real> d_qli;             // local-inertial: edge discharge, for the donor pass
    DevArray<real> d_dw_in, d_dw_tw;  // dry-weather step: node inflow and tailwater override
    int n_inlet_slots = 0;
    bool li = false;                  // local-inertial surface mode
    bool shared = false;              // shared-mesh batch: edge and cell->edge tables stored once
    DevArray<unsigned long long> d_cfl;   // CFL minimum slots [member][parity][slot] (fh::cfl_slot)
    std::vector<unsigned long long> cfl_inf;
    DevArray<real> d_nx, d_ny, d_len, d_recR, d_wet, d_area, d_n2, d_lc, d_rainm, d_f0, d_fc, d_k;
    DevArray<real> d_stage_h;   // stage ghost depth per edge position (code-3 edges with stage values)
    DevArray<unsigned short> d_cls;   // source class per cell (when use_classes)
    DevArray<V> d_cls_tab;            // class -> {n^2, m, f0, fc}
    DevArray<real> d_cls_hk;          // class -> k
    bool use_classes = false;
    DevArray<V> d_rec, d_state, d_comp;
    DevArray<double> d_part_upd, d_part_inlet, d_part_node, d_part_ret,
        d_part_ext;
    DevArray<float> d_rain_s, d_rain_e, d_rain_r, d_ext_t, d_ext_q, d_ext_ov;
    // drainage
    DevArray<real> d_inv, d_As, d_dmax, d_V, d_Vp, d_sfac, d_over, d_fb, d_wsum;
    DevArray<V> d_outfall;
    DevArray<int> d_nflag, d_ni_ptr, d_ni_cell, d_nl_ptr, d_nl_ent, d_nl2_ptr, d_nl2_ent, d_nl2_oth, d_ne_ptr;
    DevArray<int> d_ic_cells, d_ic_ptr, d_ic_node, d_ic_slot;
    DevArray<real> d_ni_surq, d_ni_net, d_ic_capq, d_ne_sw;
    DevArray<int> d_geom;
    DevArray<real> d_Q, d_req;
    DevArray<int> d_snorm;
    DevArray<int> d_rc_cells, d_rc_eptr, d_rc_enode, d_rc_fptr, d_rc_fidx, d_rc_rptr, d_rc_ridx;
    DevArray<real> d_rc_esw;
    DevArray<int> d_nle, d_nl2e, d_nl2o, d_rcf_node, d_rcf_fb;
    DevArray<int> d_lb_ptr, d_it_l, d_it_a, d_it_b, d_it_a0, d_it_a1, d_it_b0, d_it_b1, d_nl_src, d_nls;
    DevArray<int> d_ie_cell, d_ie_row, d_ie_s0, d_ie_ns, d_ie_sc;
    DevArray<real> d_ie_inv, d_ie_As, d_ie_sq, d_ic_h;
    size_t link_node_lds = sizeof(real);   // dynamic LDS of k_link_node
    size_t inlet_lds = sizeof(real);       // dynamic LDS of k_inlet
    DevArray<V> d_lk_a, d_lk_b;
    DevArray<real> d_ic_area, d_rc_area;
    DevArray<real> d_nesw0, d_rcf_sw;
    DevArray<real> d_rS, d_rcc, d_rrc, d_rstop, d_rstore;
    DevArray<int> d_rcv_qbase, d_rcv_stride, d_xc_cells, d_xc_ptr, d_xc_idx;
    fh::DrainDev<real> dd{};
    static constexpr int kTailSteps = 4;
};

struct Ctx {
    std::unique_ptr<EngineBase> eng;
    std::string err;
};

// Streaming copy, one 16-byte element per thread: the reference point for HBM
// byte counters and for the achievable bandwidth (reads and writes n floats
// once).  NT: non-temporal loads and stores.  On gfx950 this form copies 256 MiB
// faster than a grid-stride loop over 8,192 blocks (6.2 against 5.5 TB/s), and
// its non-temporal form faster still (6.7 TB/s).
template <bool NT>
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    if constexpr (NT) {
        typedef float v4 __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(__builtin_nontemporal_load((const v4*)src + i), (v4*)dst + i);
    } else {
        dst[i] = src[i];
    }
}

// ---------------------------------------------------------------------------
// Device check of fh::Divisor<float> against IEEE division (fh_division_selftest).
// c[0]/c[1]: reciprocals y != 1.0f / b, of those tested (every significand at
// each exponent of kDivExps); c[2]/c[3]: quotients != a / b over random pairs
// in the exact domain (a = +-0, or 2^-100 <= |a|, |a / b| <= 2^100, with
// 2^-100 <= b <= 2^100); c[4]/c[5]: the same for dividends below 2^-100, where
// the quotient may be one ulp off (reported, outside the domain).
// ---------------------------------------------------------------------------
constexpr int kDivExps[] = {-100, -64, -20, -1, 0, 1, 20, 64, 99};
constexpr int kDivRandBlocks = 16384, kDivRandReps = 16;

__device__ __forceinline__ bool same_bits(float x, float y) { return __float_as_uint(x) == __float_as_uint(y); }

__device__ __forceinline__ uint32_t div_hash(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// block sum of per-thread counts, one atomic add per block (vector atomics)
__device__ __forceinline__ void div_count(unsigned long long* c, unsigned v) {
    __shared__ unsigned part[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(c, (unsigned long long)t);
    }
}

__global__ void __launch_bounds__(256) k_div_recip(unsigned long long* __restrict__ c, int exp2) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;   // significand, < 2^23 by the grid
    const float b = __uint_as_float(((uint32_t)(exp2 + 127) << 23) | i);
    const fh::Divisor<float> d(b);
    div_count(c, same_bits(d.y, 1.0f / b) ? 0u : 1u);
}

__global__ void __launch_bounds__(256) k_div_pairs(unsigned long long* __restrict__ c, uint32_t seed) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    unsigned bad = 0, n = 0, bad_tiny = 0, n_tiny = 0;
    for (int k = 0; k < kDivRandReps; ++k) {
        const uint32_t h1 = div_hash(t * 2654435761u + seed + (uint32_t)k * 0x9e3779b9u);
        const uint32_t h2 = div_hash(h1 ^ 0x85ebca6bu), h3 = div_hash(h2 + 0x27d4eb2fu);
        const int eb = -100 + (int)(h3 % 201u);
        const float b = __uint_as_float(((uint32_t)(eb + 127) << 23) | (h2 >> 9));
        const uint32_t sign = (h1 & 1u) << 31;
        const uint32_t sel = (h3 >> 16) & 63u;
        float a;
        if (sel == 0) {
            a = __uint_as_float(sign);                                   // +-0
        } else if (sel < 8) {
            const int ea = -126 + (int)((h3 >> 8) % 26u);                // below 2^-100
            a = __uint_as_float(sign | ((uint32_t)(ea + 127) << 23) | (h1 >> 9));
        } else {
            const int ea = -100 + (int)((h3 >> 8) % 201u);
            a = __uint_as_float(sign | ((uint32_t)(ea + 127) << 23) | (h1 >> 9));
        }
        const float ref = a / b;
        const bool ok = same_bits(fh::Divisor<float>(b).div(a), ref);
        const float aa = fabsf(a), ar = fabsf(ref);
        if (aa != 0.0f && aa < 0x1p-100f) {
            ++n_tiny;
            bad_tiny += ok ? 0u : 1u;
        } else if (aa == 0.0f || (ar >= 0x1p-100f && ar <= 0x1p100f && aa <= 0x1p100f)) {
            ++n;
            bad += ok ? 0u : 1u;
        }
    }
    div_count(c + 2, bad);
    __syncthreads();
    div_count(c + 3, n);
    __syncthreads();
    div_count(c + 4, bad_tiny);
    __syncthreads();
    div_count(c + 5, n_tiny);
}

// Device check of fh::fric_pow<float> (fh_pow_selftest): every float x in
// [2^kPowExp0, 2^kPowExp1) against the correctly rounded x^p, p = (float)(4/3),
// taken as the FP64 pow rounded to float; the device library's powf beside it.
// c[0] fric_pow mismatches, c[1] values, c[2] powf mismatches, c[3] and c[4]
// the largest ulp distance of each.
constexpr int kPowExp0 = -40, kPowExp1 = 20;

__global__ void __launch_bounds__(256) k_pow_check(unsigned long long* __restrict__ c, double dln2) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;   // < (kPowExp1 - kPowExp0) << 23 by the grid
    const float x = __uint_as_float((uint32_t)(kPowExp0 + 127) * (1u << 23) + i);
    const float p = (float)(4.0 / 3.0);
    const float ref = (float)pow((double)x, (double)p);
    const float mine = fh::fric_pow(x, dln2);
    const float lib = powf(x, p);
    const int dm = abs((int)__float_as_uint(mine) - (int)__float_as_uint(ref));
    const int dl = abs((int)__float_as_uint(lib) - (int)__float_as_uint(ref));
    div_count(c, dm ? 1u : 0u);
    __syncthreads();
    div_count(c + 2, dl ? 1u : 0u);
    if (dm) atomicMax(&c[3], (unsigned long long)dm);
    if (dl) atomicMax(&c[4], (unsigned long long)dl);
}

}  // namespace

struct fh_ctx {
    Ctx c;
};

// ---------------------------------------------------------------------------
// Allocation-free CFL minimum (fh_stable_timestep_noalloc).  The same per-cell
// CFL time and minimum as fh_stable_timestep, on caller-owned device arrays; the
// per-block minima go to `scratch`, a file-scope device array (each GPU the
// library runs on holds its own copy until the process ends), instead of the
// context's CFL slots.  Host code never passes the array's name to a kernel: it
// looks the array's device address up once per device (scratch_address) and
// passes that.
// ---------------------------------------------------------------------------
constexpr int kNoAllocBlocks = 1024;     // block minima: 8 KiB per device
constexpr int kNoAllocThreads = 256;
constexpr int kNoAllocCellsPerBlock = 2048;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocBlocks];

std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first
// call there (hipGetSymbolAddress, with that device current); nullptr and the
// HIP status in *err on failure.
double* scratch_address(int device, hipError_t* err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) { *err = e; return nullptr; }
        g_scratch_addr[device] = static_cast<double*>(p);
    }
    return g_scratch_addr[device];
}

// Block minimum of the CFL time over a grid-stride range of cells into part[block].
__global__ void __launch_bounds__(kNoAllocThreads) k_cfl_noalloc(const float* __restrict__ h,
        const float* __restrict__ hu, const float* __restrict__ hv, const float* __restrict__ lc, int64_t n,
        fh::KP<float> p, double* __restrict__ part) {
    __shared__ double lds[16];
    float mn = INFINITY;
    for (int64_t c = (int64_t)blockIdx.x * kNoAllocThreads + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * kNoAllocThreads) {
        const float4 st = make_float4(h[c], hu[c], hv[c], 0.0f);
        const float local = fh::cell_cfl<float>(p, st, lc[c]);
        mn = local < mn ? local : mn;
    }
    const double v = fh::block_min((double)mn, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Minimum of the nparts block minima (all written by the launch before on the
// same stream) into *out; +inf when there are none.
__global__ void __launch_bounds__(kNoAllocThreads) k_cfl_noalloc_final(const double* __restrict__ part,
                                                                      int nparts, double* __restrict__ out) {
    __shared__ double lds[16];
    double mn = INFINITY;
    for (int i = threadIdx.x; i < nparts; i += kNoAllocThreads) {
        const double x = part[i];
        mn = x < mn ? x : mn;
    }
    mn = fh::block_min(mn, lds);
    if (threadIdx.x == 0) *out = mn;
}

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
    p->block_size = 64;
    p->xcd_remap = 1;
    p->device = 0;
    p->edges_per_thread = 1;
    p->pack_order = 0;
    p->cell_order = 1;
    p->source_classes = 1;
    p->surface_mode = 0;
    p->share_mesh = 1;
}

const char* fh_version(void) { return "floodhip 0.2.0 (gfx950)"; }

int fh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* fh_global_error(void) { return g_error.c_str(); }

fh_ctx* fh_create_batch(const fh_params* params, int32_t members) {
    fh_params p;
    if (params) p = *params; else fh_default_params(&p);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_error = "no HIP device is visible (BackendUnavailableError)";
        return nullptr;
    }
    if (p.device < 0 || p.device >= n) { g_error = "device ordinal out of range"; return nullptr; }
    if (p.precision != 32 && p.precision != 64) { g_error = "precision must be 32 or 64"; return nullptr; }
    if (members < 1 || members > 65536) { g_error = "members must be in [1, 65536]"; return nullptr; }
    try {
        auto* ctx = new fh_ctx();
        if (p.precision == 64) ctx->c.eng.reset(new Engine<double>(p, members));
        else ctx->c.eng.reset(new Engine<float>(p, members));
        return ctx;
    } catch (const HipError& h) {
        g_error = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.where;
    } catch (const std::exception& x) {
        g_error = x.what();
    }
    return nullptr;
}

fh_ctx* fh_create(const fh_params* params) { return fh_create_batch(params, 1); }

void fh_destroy(fh_ctx* ctx) { delete ctx; }

const char* fh_last_error(const fh_ctx* ctx) { return ctx ? ctx->c.err.c_str() : g_error.c_str(); }

int fh_select_member(fh_ctx* ctx, int32_t member) {
    FH_GUARD(ctx, ctx->c.eng->select(member));
    return FH_OK;
}

int32_t fh_member_count(const fh_ctx* ctx) { return ctx ? ctx->c.eng->members() : 0; }

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

int fh_set_boundary_stage(fh_ctx* ctx, int64_t n_values, const float* stage_m) {
    FH_GUARD(ctx, ctx->c.eng->set_boundary_stage(n_values, stage_m));
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

int fh_get_drainage_f64(fh_ctx* ctx, double* node_volume, double* link_flow) {
    FH_GUARD(ctx, ctx->c.eng->get_drainage_f64(node_volume, link_flow));
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

int fh_dry_weather_step(fh_ctx* ctx, double dt, const float* node_inflow_m3ps, const float* outfall_tailwater_head_m,
                        int32_t cap_node_overflow, double* diag_out) {
    if (!diag_out) return FH_ERR_VALUE;
    FH_GUARD(ctx, ctx->c.eng->dry_weather_step(dt, node_inflow_m3ps, outfall_tailwater_head_m, cap_node_overflow,
                                               diag_out));
    return FH_OK;
}

int fh_stable_timestep(fh_ctx* ctx, double* dt_out) {
    if (!dt_out) return FH_ERR_VALUE;
    FH_GUARD(ctx, *dt_out = ctx->c.eng->stable_timestep());
    return FH_OK;
}

int fh_stable_timestep_noalloc(const float* h, const float* hu, const float* hv,
                               const float* characteristic_length, int64_t n_cells, double cfl,
                               double* dt_out, void* stream) {
    if (n_cells < 0 || n_cells > FH_NOALLOC_MAX_CELLS) {
        g_error = "fh_stable_timestep_noalloc: n_cells must be in [0, 2^31 - 1]";
        return FH_ERR_VALUE;
    }
    if (!dt_out || (n_cells > 0 && (!h || !hu || !hv || !characteristic_length))) {
        g_error = "fh_stable_timestep_noalloc: NULL device array";
        return FH_ERR_VALUE;
    }
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess || device < 0 || device >= kMaxDevices) {
        g_error = std::string("fh_stable_timestep_noalloc: no usable current device: ") + hipGetErrorString(e);
        return FH_ERR_UNAVAILABLE;
    }
    double* part = scratch_address(device, &e);
    if (!part) {
        g_error = std::string("fh_stable_timestep_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
        return FH_ERR_RUNTIME;
    }
    fh::KP<float> p{};
    p.g = (float)9.80665;
    p.eps = (float)1.0e-5;
    p.cfl = (float)cfl;
    p.tiny = (float)1.0e-12;
    const int64_t want = (n_cells + kNoAllocCellsPerBlock - 1) / kNoAllocCellsPerBlock;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(kNoAllocBlocks, want));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    k_cfl_noalloc<<<blocks, kNoAllocThreads, 0, s>>>(h, hu, hv, characteristic_length, n_cells, p, part);
    k_cfl_noalloc_final<<<1, kNoAllocThreads, 0, s>>>(part, blocks, dt_out);
    e = hipGetLastError();
    if (e != hipSuccess) {
        g_error = std::string("fh_stable_timestep_noalloc: launch failed: ") + hipGetErrorString(e);
        return FH_ERR_RUNTIME;
    }
    return FH_OK;
}

int fh_pow_selftest(int64_t* counts) {
    if (!counts) return FH_ERR_VALUE;
    unsigned long long* c = nullptr;
    int status = FH_OK;
    if (hipMalloc((void**)&c, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(c, 0, 8 * sizeof(unsigned long long)) != hipSuccess) {
        status = FH_ERR_RUNTIME;
    } else {
        const unsigned long long n = (unsigned long long)(kPowExp1 - kPowExp0) << 23;
        k_pow_check<<<(unsigned)(n / 256), 256>>>(c, ((double)(float)(4.0 / 3.0) - 4.0 / 3.0) * M_LN2);
        unsigned long long h[8] = {};
        if (hipGetLastError() != hipSuccess || hipMemcpy(h, c, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
            status = FH_ERR_RUNTIME;
        } else {
            h[1] = n;
            for (int k = 0; k < 5; ++k) counts[k] = (int64_t)h[k];
        }
    }
    if (c) (void)hipFree(c);
    if (status != FH_OK) g_error = "pow self-test failed to run";
    return status;
}

int fh_division_selftest(int64_t* counts) {
    if (!counts) return FH_ERR_VALUE;
    unsigned long long* c = nullptr;
    int status = FH_OK;
    if (hipMalloc((void**)&c, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(c, 0, 8 * sizeof(unsigned long long)) != hipSuccess) {
        status = FH_ERR_RUNTIME;
    } else {
        for (int e : kDivExps) k_div_recip<<<(1 << 23) / 256, 256>>>(c, e);
        k_div_pairs<<<kDivRandBlocks, 256>>>(c, 20261204u);
        unsigned long long h[8] = {};
        if (hipGetLastError() != hipSuccess || hipMemcpy(h, c, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
            status = FH_ERR_RUNTIME;
        } else {
            h[1] = (unsigned long long)(sizeof(kDivExps) / sizeof(kDivExps[0])) << 23;
            for (int k = 0; k < 6; ++k) counts[k] = (int64_t)h[k];
        }
    }
    if (c) (void)hipFree(c);
    if (status != FH_OK) g_error = "division self-test failed to run";
    return status;
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

int fh_copy_probe(int64_t n_floats, int32_t reps, int32_t nontemporal, double* avg_ms) {
    if (n_floats <= 0 || (n_floats & 3) || n_floats / 4 > (int64_t)INT32_MAX * 256 || reps <= 0 || !avg_ms)
        return FH_ERR_VALUE;
    float* a = nullptr;
    float* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = FH_OK;
    if (hipMalloc((void**)&a, n_floats * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&b, n_floats * sizeof(float)) != hipSuccess ||
        hipMemset(a, 0, n_floats * sizeof(float)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        status = FH_ERR_RUNTIME;
    } else {
        const int64_t n4 = n_floats / 4;
        const int grid = (int)((n4 + 255) / 256);
        auto launch = [&]() {
            if (nontemporal) k_copy_probe<true><<<grid, 256>>>((const float4*)a, (float4*)b, n4);
            else k_copy_probe<false><<<grid, 256>>>((const float4*)a, (float4*)b, n4);
        };
        launch();   // warm up
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            status = FH_ERR_RUNTIME;
        *avg_ms = (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (status != FH_OK) g_error = "copy probe failed";
    return status;
}

}  // extern "C"
