// The engine: per-size device workspaces, the pair pipeline (coarse stage, fine stage,
// confidence, guard, strain) on one HIP stream per lane, the chained sequence, and the C ABI of
// include/ofx.h.
//
// Reference call structure: rust/crates/optflow-core/src/track.rs:567-861 (pair),
// rust/crates/optflow-core/src/phaseflow.rs:1514-1613 (fine stage loop),
// rust/crates/optflow-core/src/track.rs:1112-1225 (chained sequence),
// rust/crates/optflow-gpu/src/engine.rs:633-1071 (lanes, check_window, errors).
#include <atomic>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "cnn.hpp"
#include "fine.hpp"

namespace ofx {

// Largest side of an image-2 window: a row of its spline prefilter is held in LDS (two padded
// float lines within 64 KiB).
constexpr int MAX_I2_SIDE = 7680;

// ---------------------------------------------------------------- watchdog
// A fixed pool of per-thread progress records; a thread takes a free one on its first call and
// returns it when it exits. One detached thread reports on stderr, once per call and thread, every
// thread whose call has run past the limit. It only reads the atomics, never calls HIP, so a
// thread blocked inside the runtime cannot block it.
namespace {
constexpr int NPROG = 256;
Progress g_prog[NPROG];
Progress g_overflow;  // shared by threads beyond the pool (reports then name the last writer)

struct ProgressHolder {
    Progress *p = nullptr;
    ProgressHolder() {
        for (int i = 0; i < NPROG && !p; i++) {
            bool expected = false;
            if (g_prog[i].used.compare_exchange_strong(expected, true)) p = &g_prog[i];
        }
        if (!p) p = &g_overflow;
        p->step = "-";
        p->hip_call = nullptr;
        p->call_since_ms = 0;
        p->pending = -1;
    }
    ~ProgressHolder() {
        if (p != &g_overflow) p->used.store(false);
    }
};

void watchdog_start() {
    static std::once_flag once;
    std::call_once(once, [] {
        const char *v = std::getenv("OFX_WATCHDOG_S");
        const double limit_s = v ? std::atof(v) : 60.0;
        if (limit_s <= 0) return;
        std::thread([limit_s] {
            static long long reported[NPROG + 1] = {0};
            for (;;) {
                std::this_thread::sleep_for(std::chrono::seconds(1));
                const long long now = now_ms();
                for (int i = 0; i <= NPROG; i++) {
                    Progress &p = i < NPROG ? g_prog[i] : g_overflow;
                    const long long t0 = p.call_since_ms.load(), id = p.call_id.load();
                    if (t0 == 0 || reported[i] == id || (now - t0) * 1e-3 < limit_s) continue;
                    const char *call = p.hip_call.load();
                    const long long hs = p.hip_since_ms.load();
                    const int pend = p.pending.load();
                    std::fprintf(stderr,
                                 "ofx watchdog: thread slot %d: call running %.0f s; stage %s; %s%s%s%.0f s; last "
                                 "kernel launch %s:%d; last wait found device work %s\n",
                                 i, (now - t0) * 1e-3, p.step.load(), call ? "inside " : "in no HIP call (", call ? call : "",
                                 call ? " for " : "host code for ", call ? (now - hs) * 1e-3 : 0.0, p.launch_file.load(),
                                 p.launch_line.load(), pend < 0 ? "(no wait yet)" : pend ? "pending" : "done");
                    std::fflush(stderr);
                    reported[i] = id;
                }
            }
        }).detach();
    });
}
}  // namespace

Progress &progress() {
    static thread_local ProgressHolder h;
    return *h.p;
}
long long now_ms() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
void step_mark(const char *what) { progress().step.store(what, std::memory_order_relaxed); }
void launch_mark(const char *file, int line) {
    Progress &p = progress();
    p.launch_file.store(file, std::memory_order_relaxed);
    p.launch_line.store(line, std::memory_order_relaxed);
}
HipCallMark::HipCallMark(const char *call) {
    Progress &p = progress();
    p.hip_since_ms.store(now_ms(), std::memory_order_relaxed);
    p.hip_call.store(call, std::memory_order_relaxed);
}
HipCallMark::~HipCallMark() { progress().hip_call.store(nullptr, std::memory_order_relaxed); }

namespace {
template <class Q>
void poll(Q query, const char *what) {
    Progress &p = progress();
    HipCallMark mark(what);
    for (long k = 0;; k++) {
        const hipError_t e = query();
        if (e == hipSuccess) {
            p.pending.store(0, std::memory_order_relaxed);
            return;
        }
        if (e != hipErrorNotReady) throw Error(OFX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        p.pending.store(1, std::memory_order_relaxed);
        if (k < 4000)
            std::this_thread::yield();
        else
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}
}  // namespace

void wait_stream(hipStream_t s) { poll([&] { return hipStreamQuery(s); }, "hipStreamQuery (waiting for a stream)"); }
void wait_event(hipEvent_t e) { poll([&] { return hipEventQuery(e); }, "hipEventQuery (waiting for an event)"); }

namespace {
thread_local DeferFrees *g_defer = nullptr;
}
DeferFrees::DeferFrees() : prev(g_defer) { g_defer = this; }
DeferFrees::~DeferFrees() {
    g_defer = prev;
    for (void *p : held) (void)hipFree(p);
}
void release_device(void *p) {
    if (!p) return;
    if (g_defer)
        g_defer->held.push_back(p);
    else
        (void)hipFree(p);
}

PinnedBuf::~PinnedBuf() {
    if (p) (void)hipHostFree(p);
}
char *PinnedBuf::get() {
    if (!p) {
        HIPCHECK(hipHostMalloc((void **)&p, PIN_CHUNK, hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHECK(hipHostGetDevicePointer((void **)&dp, p, 0));
    }
    return p;
}
void copy_h2d(PinnedBuf &pin, void *dev, const void *host, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    char *b = pin.get();
    for (size_t off = 0; off < bytes; off += PIN_CHUNK) {
        const size_t len = std::min(PIN_CHUNK, bytes - off);
        wait_stream(s);  // the staging buffer is free again
        std::memcpy(b, (const char *)host + off, len);
        dcopy((char *)dev + off, pin.dp, len, s);
    }
    wait_stream(s);
}
void copy_d2h(PinnedBuf &pin, void *host, const void *dev, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    char *b = pin.get();
    for (size_t off = 0; off < bytes; off += PIN_CHUNK) {
        const size_t len = std::min(PIN_CHUNK, bytes - off);
        dcopy(pin.dp, (const char *)dev + off, len, s);
        wait_stream(s);
        std::memcpy((char *)host + off, b, len);
    }
}

namespace {

thread_local std::string g_last_error;

std::vector<double> gauss_taps(double sigma) {
    const int r = (int)(4.0 * sigma + 0.5);
    std::vector<double> k(2 * r + 1);
    double s = 0;
    for (int i = -r; i <= r; i++) {
        const double t = i / sigma;
        k[i + r] = std::exp(-0.5 * t * t);
        s += k[i + r];
    }
    for (auto &v : k) v /= s;
    return k;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T *get(size_t n) {
        const size_t b = n * sizeof(T);
        if (bytes < b) {
            if (p) release_device(p);
            p = nullptr;
            bytes = 0;
            HIPCHECK(hipMalloc(&p, b));
            bytes = b;
        }
        return (T *)p;
    }
};

}  // namespace

// Per-stage GPU time (the stages of rust/crates/optflow-core/src/prof.rs:17-44): an event marks
// the start of every stage on the lane's stream; at the end of a call the time between
// consecutive marks is added to the earlier mark's stage.
enum Stage {
    ST_COARSE, ST_REFERENCE, ST_PREWARP, ST_SMOOTH, ST_WARP, ST_RESPONSES, ST_ACCUMULATE, ST_MOMENTS, ST_SOLVE,
    ST_CONFIDENCE, ST_GUARD, ST_STRAIN, ST_END, ST_COUNT
};
const char *const STAGE_NAMES[ST_COUNT] = {"coarse", "reference", "prewarp", "smooth", "warp", "responses",
                                           "accumulate", "moments", "solve", "confidence", "guard", "strain", "end"};

struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<int> stage;
    size_t used = 0;
    double ms[ST_COUNT] = {0};
    long long calls[ST_COUNT] = {0};
    void mark(int st, hipStream_t s) {
        if (!on) return;
        if (used == pool.size()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        HIPCHECK(hipEventRecord(pool[used], s));
        if (stage.size() < pool.size()) stage.resize(pool.size());
        stage[used++] = st;
    }
    void flush() {
        if (!on || used == 0) return;
        wait_event(pool[used - 1]);
        for (size_t i = 0; i + 1 < used; i++) {
            float t = 0;
            HIPCHECK(hipEventElapsedTime(&t, pool[i], pool[i + 1]));
            ms[stage[i]] += t;
            calls[stage[i]] += 1;
        }
        used = 0;
    }
    void reset() {
        used = 0;
        for (int i = 0; i < ST_COUNT; i++) {
            ms[i] = 0;
            calls[i] = 0;
        }
    }
    ~Profiler() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Device buffers of one lane, grown to the largest window seen (release_caches frees them).
struct Workspace {
    Profiler prof;
    int w = 0, h = 0, F = 0;
    DevBuf bi1, bi2, bmask, binx, biny;                 // staging of host inputs
    DevBuf bc2, bspec, btmp, btmp2, bresp, brefp, bpack;
    // the reference pack of a sequence's reference frame (refine_device, cache_i1)
    bool pack_valid = false;
    int pack_w = 0, pack_h = 0;
    Bank pack_bank{};
    DevBuf bux, buy, bnux, bnuy, bsx, bsy, bcsx, bcsy, bwarped, bfx, bfy, bbt;
    DevBuf bsums, bym, benergy, bresid, bslope, bmed, bmedscr, bmom2;
    DevBuf bconf, bconf64, blow, bkept, bwithin, bok, bcold2, bcolrow, bcnt;
    DevBuf bparent, bsizes, blarge, bsrc, bnear, bmx, bmy, bytmp, bcolsum, bpftmp;
    DevBuf bstart_x, bstart_y;
    DevBuf btav_acc, btav_a, btav_fx, btav_fy, btav_c, btav_tmp, bpix;  // temporal averaging, frame input
    DevBuf bfsx, bfsy, bjx, bjy, bje, bjr, btile1, btile2, bpart;         // tiled fine stage
    PinnedBuf pin;                                                       // staging of host copies
    DevBuf bout;                                       // staging of host outputs (f32 x 10 + u8)
    DevBuf btaps_s, btaps_w;
    double taps_s_sigma = -1, taps_w_sigma = -1;
    std::vector<double> taps_s_host, taps_w_host;
    int taps_s_r = 0, taps_w_r = 0;
    // sequence state
    DevBuf bu0x, bu0y, bc0r, bcx, bcy, bconf0, bca, bprev_x, bprev_y, bprev_c, bleg_c, bmean, bframe_prev, bframe_ref;
    // hybrid sequence state (track/hybrid.rs)
    DevBuf banch, bgood_prev, bupx, bupy, bnfb, bpadj_x, bpadj_y, badj_x, badj_y, badj_c, blink, bflags, bax, bay,
        bcax, bcay, bca_adj, bgate0, bgate, bhcnt, bgscr, bhlow, bhux, bhuy, bfb;
    CoarseWs *coarse = nullptr;
    // captured fine stages (refine_graph): the keys seen once eagerly, and two graphs
    std::vector<std::vector<uint64_t>> warm_keys;
    struct FineGraph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    } graphs[2];
    int graph_next = 0;
    bool graphs_off = false;  // a capture failed: this workspace runs eagerly from then on
    // the engine's schedule (ofx_set_schedule), copied in when a call takes the lane
    int tile_groups = 3;
    bool graphs_on = true;
    long long graph_launches = 0, graph_captures = 0;
    // the engine's learned coarse network (read only during a call) and this lane's activations
    const CoarseNet *net = nullptr;
    std::unique_ptr<NetActs> nacts;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // the tiled fine stage: the workspaces and streams of the tile groups that run beside the
    // lane's own stream, and the events that fork them from it and join them back (first use)
    std::vector<std::unique_ptr<Workspace>> sub;
    std::vector<hipStream_t> sub_streams;
    std::vector<hipEvent_t> sub_done;
    hipEvent_t fork = nullptr;
    ~Workspace() {
        for (auto &g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        coarse_release(coarse);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t t : sub_streams) {
            (void)hipStreamSynchronize(t);
            (void)hipStreamDestroy(t);
        }
        for (hipEvent_t e : sub_done) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
    }
};

struct Lane {
    hipStream_t stream = nullptr;
    std::unique_ptr<Workspace> ws;
    std::atomic<bool> busy{false};
};

}  // namespace ofx

struct ofx_engine {
    int device = 0;
    std::vector<std::unique_ptr<ofx::Lane>> lanes;
    std::unique_ptr<ofx::CoarseNet> net;  // the learned coarse network (ofx_coarse_net_load)
    std::atomic<int> tile_groups{3};      // ofx_set_schedule
    std::atomic<bool> graphs{true};
};

namespace ofx {

namespace {

struct LaneGuard {
    Lane *lane = nullptr;
    explicit LaneGuard(ofx_engine *e) {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        for (auto &l : e->lanes) {
            bool expected = false;
            if (l->busy.compare_exchange_strong(expected, true)) {
                lane = l.get();
                break;
            }
        }
        if (!lane) throw Error(OFX_EBUSY, "every lane of the GPU engine is busy");
        HIPCHECK(hipSetDevice(e->device));
        if (!lane->ws) lane->ws.reset(new Workspace);
        lane->ws->net = e->net.get();
        lane->ws->tile_groups = e->tile_groups.load();
        lane->ws->graphs_on = e->graphs.load();
    }
    ~LaneGuard() {
        if (lane) lane->busy.store(false);
    }
};

double max_scale(const ofx_config &c) {
    double m = c.scales[0];
    for (int i = 1; i < c.n_scales; i++) m = std::max(m, c.scales[i]);
    return m;
}
double auto_edge_guard(const ofx_config &c) { return 3.0 * c.lam * max_scale(c) / (2 * M_PI * c.bandwidth); }
double edge_guard(const ofx_config &c) { return c.edge_guard >= 0 ? c.edge_guard : auto_edge_guard(c); }
// config.py:253-255: half-width of the hybrid gate's window
double fb_radius_px(const ofx_config &c) { return c.fb_radius > 0 ? c.fb_radius : 4.0 * c.sigma_w; }
double edge_guard_min_area(const ofx_config &c) {
    if (c.edge_guard_min_area >= 0) return c.edge_guard_min_area;
    const double l = c.lam * max_scale(c);
    return 2.0 * l * l;
}

// The reference's checks (python/optflow_ref/config.py:166-245) for the keys the engine reads.
void validate(const ofx_config *cfg) {
    if (!cfg) throw Error(OFX_EINVAL, "config is NULL");
    const ofx_config &c = *cfg;
    if (c.n_scales < 1 || c.n_scales > OFX_MAX_SCALES) throw Error(OFX_EINVAL, "scales must not be empty");
    if (c.poc_win < 16 || (c.poc_win & (c.poc_win - 1))) throw Error(OFX_EINVAL, "poc_win must be a power of two >= 16");
    if (c.grid_step < 1 || c.strain_win < 5 || c.strain_win % 2 == 0)
        throw Error(OFX_EINVAL, "grid_step >= 1 and strain_win odd >= 5 required");
    if (!(c.dc_cut > 0)) throw Error(OFX_EINVAL, "dc_cut must be positive");
    if (!(c.phase_tol > 0)) throw Error(OFX_EINVAL, "phase_tol must be positive");
    if (c.shape_order != 1 && c.shape_order != 2) throw Error(OFX_EINVAL, "shape_order must be 1 or 2");
    double mins = c.scales[0];
    for (int i = 1; i < c.n_scales; i++) mins = std::min(mins, c.scales[i]);
    if (c.lam * mins < 3 || !(c.sigma_w > 0) || !(c.sigma_s > 0))
        throw Error(OFX_EINVAL, "shortest wavelength lam*min(scales) must be >= 3 px; sigmas must be positive");
    if (!(c.bandwidth > 0) || !(c.max_disp > 0) || !(c.median_thresh > 0))
        throw Error(OFX_EINVAL, "bandwidth, max_disp and median_thresh must be positive");
    if (c.n_orient < 1 || c.outer_iters < 1 || c.inner_iters < 1)
        throw Error(OFX_EINVAL, "n_orient, outer_iters and inner_iters must be >= 1");
    if (!(c.conf_min >= 0 && c.conf_min <= 1) || !(c.poc_peak_min >= 0 && c.poc_peak_min <= 1))
        throw Error(OFX_EINVAL, "conf_min and poc_peak_min must lie in [0, 1]");
    if (c.prewarp_freeze_at < 0) throw Error(OFX_EINVAL, "prewarp_freeze_at must be >= 0");
    if ((c.coarse_learned != 0 && c.coarse_learned != 1) || !(c.coarse_gate >= 0 && c.coarse_gate <= 1))
        throw Error(OFX_EINVAL, "coarse_learned must be 0 or 1 and coarse_gate must lie in [0, 1]");
    if (c.tile != 0 && c.tile < 64) throw Error(OFX_EINVAL, "tile must be 0 (whole frame) or an integer >= 64");
    if (c.n_orient * c.n_scales > 64 || c.sigma_w > 256 || c.sigma_s > 256 || c.strain_win > 255 || c.poc_win > 1024)
        throw Error(OFX_EINVAL,
                    "at most 64 filters (n_orient * len(scales)); sigma_w, sigma_s <= 256, strain_win <= 255 and "
                    "poc_win <= 1024 required");
}

// tile.rs:80-95
int tile_margin(const ofx_config &c) {
    return std::max({64, (int)std::ceil(4.0 * c.sigma_w), (int)std::ceil(4.0 * c.sigma_s),
                     (int)std::ceil(std::max(edge_guard(c), 0.0)), c.strain_win});
}

// The largest window the fine stage refines for a w x h frame must fit the line transforms (one
// padded line per workgroup's LDS): the whole frame, or with tiling (c.tile > 0) a tile's
// extension, at most c.tile + 2 * tile_margin a side (the reference checks
// check_window(min(w, side), min(h, side)), track.rs:754-762). The other stages have no line
// limit; a tile's window of image 2 is checked against MAX_I2_SIDE when the tiles are laid out.
void check_window_impl(int w, int h, const ofx_config &c) {
    if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty window");
    if (w <= PAD || h <= PAD) throw Error(OFX_EINVAL, "image smaller than the 32 px reflect pad");
    int fw = w, fh = h;
    if (c.tile > 0) {
        const int side = c.tile + 2 * tile_margin(c);
        fw = std::min(w, side);
        fh = std::min(h, side);
    }
    if (fw + 2 * PAD > 4096 || fh + 2 * PAD > 4096)
        throw Error(OFX_ETOOLARGE, "window " + std::to_string(fw) + "x" + std::to_string(fh) +
                                       " exceeds 4032 px per side (one padded line must fit a workgroup's LDS)" +
                                       (c.tile > 0 ? "" : "; a tile size refines larger frames tile by tile"));
}

template <class T>
T *B(DevBuf &b, size_t n) {
    return b.get<T>(n);
}

// the checks of track.py:270-279 (and track.rs:655-663) before a sequence starts
void check_sequence_args(int w, int h, const uint8_t *mask, const ofx_config &c) {
    check_window_impl(w, h, c);
    if (std::min(w, h) < c.poc_win)
        throw Error(OFX_EINVAL, "image smaller than poc_win (" + std::to_string(c.poc_win) + ")");
    if (mask) {
        const size_t n = (size_t)w * h;
        bool any = false;
        for (size_t i = 0; i < n && !any; i++) any = mask[i] != 0;
        if (!any) throw Error(OFX_EINVAL, "ROI mask is empty");
    }
}

void set_taps(Workspace &ws, const ofx_config &c, hipStream_t s) {
    if (ws.taps_s_sigma != c.sigma_s) {
        std::vector<double> k = gauss_taps(c.sigma_s);
        double *d = ws.btaps_s.get<double>(k.size());
        copy_h2d(ws.pin, d, k.data(), k.size() * sizeof(double), s);
        ws.taps_s_host = k;
        ws.taps_s_sigma = c.sigma_s;
        ws.taps_s_r = (int)k.size() / 2;
    }
    if (ws.taps_w_sigma != c.sigma_w) {
        std::vector<double> k = gauss_taps(c.sigma_w);
        const int r = (int)k.size() / 2, klen = 2 * r + 1;
        // rows p = 0 .. 4: the taps times t^p (order 1 uses the first three)
        std::vector<double> kt(5 * klen);
        for (int p = 0; p < 5; p++)
            for (int t = -r; t <= r; t++) kt[p * klen + t + r] = k[t + r] * std::pow((double)t, p);
        double *d = ws.btaps_w.get<double>(kt.size());
        copy_h2d(ws.pin, d, kt.data(), kt.size() * sizeof(double), s);
        ws.taps_w_host = kt;
        ws.taps_w_sigma = c.sigma_w;
        ws.taps_w_r = r;
    }
}

// The fine stage (phaseflow.py:427-511, phaseflow.rs:1441-1613) on image 1 (w x h) against image 2
// (w2 x h2), pixel (x, y) of image 1 sitting at (x + shx, y + shy) in image 2 before the field
// displaces it. The start field is in ws.bux / ws.buy (double); the result replaces it, and energy,
// resid, the energy median (ws.bmed[0]) and the final confidence (ws.bconf as float, ws.bconf64 as
// double) are left in the workspace.
bool same_bank(const Bank &a, const Bank &b) {
    if (a.F != b.F || a.s_hp2_half != b.s_hp2_half) return false;
    for (int k = 0; k < a.F; k++)
        if (a.w0x[k] != b.w0x[k] || a.w0y[k] != b.w0y[k] || a.inv2s2[k] != b.inv2s2[k]) return false;
    return true;
}

void refine_device(Workspace &ws, const float *i1, const float *i2, int w, int h, int w2, int h2, long shx, long shy,
                   const ofx_config &c, hipStream_t s, bool cache_i1) {
    const int wp = w + 2 * PAD, hp = h + 2 * PAD;
    const size_t n = (size_t)w * h, np = (size_t)wp * hp;
    const Bank bank = make_bank(c);
    const int F = bank.F;
    set_taps(ws, c, s);
    // F planes of scratch for the batched inverses; 3F response planes for the reference pack
    float2 *spec = B<float2>(ws.bspec, np), *tmp = B<float2>(ws.btmp, F * np), *tmp2 = B<float2>(ws.btmp2, F * np);
    float2 *resp = B<float2>(ws.bresp, 3 * F * np);
    // A sequence's reference frame is image 1 of the direct leg at every step until the reference
    // moves: its pack is kept in ws.bpack and computed again only after the sequence marks it
    // stale (ws.pack_valid), or when the frame size or the filter bank changed.
    const bool cached = cache_i1 && ws.pack_valid && ws.pack_w == w && ws.pack_h == h && same_bank(ws.pack_bank, bank);
    float4 *refp = cache_i1 ? B<float4>(ws.bpack, F * n) : B<float4>(ws.brefp, F * n);
    step_mark("fine: reference pack");
    ws.prof.mark(ST_REFERENCE, s);
    if (!cached) {
        // reference pack of image 1 (phaseflow.py:100-111): three inverses per filter, each kind
        // batched over the filters
        fft_forward_padded(i1, w, h, spec, tmp, tmp2, s);
        fft_inverse_filtered(spec, hp, wp, bank, 0, F, MUL_G, false, tmp, tmp2, resp, s);
        fft_inverse_filtered(spec, hp, wp, bank, 0, F, MUL_G_IFX, false, tmp, tmp2, resp + F * np, s);
        fft_inverse_filtered(spec, hp, wp, bank, 0, F, MUL_G_IFY, false, tmp, tmp2, resp + 2 * F * np, s);
        refpack(resp, F, refp, w, h, wp, s);
        if (cache_i1) {
            ws.pack_valid = true;
            ws.pack_w = w;
            ws.pack_h = h;
            ws.pack_bank = bank;
        }
    }
    // spline of image 2
    ws.prof.mark(ST_WARP, s);
    const size_t n2 = (size_t)w2 * h2;
    float *c2 = B<float>(ws.bc2, n2);
    dcopy(c2, i2, n2 * sizeof(float), s);
    // sized for the two pre-warp planes of the outer iterations too (prefilter_real2)
    prefilter_real(c2, B<float>(ws.bpftmp, std::max(2 * n, n2)), w2, h2, s);
    double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
    double *nux = B<double>(ws.bnux, n), *nuy = B<double>(ws.bnuy, n);
    float *sx = B<float>(ws.bsx, n), *sy = B<float>(ws.bsy, n), *csx = B<float>(ws.bcsx, n), *csy = B<float>(ws.bcsy, n);
    float *warped = B<float>(ws.bwarped, n);
    double *fx = B<double>(ws.bfx, n), *fy = B<double>(ws.bfy, n), *bt = B<double>(ws.bbt, n);
    double *sums = B<double>(ws.bsums, 7 * n), *ym = B<double>(ws.bym, 15 * n);
    double *energy = B<double>(ws.benergy, n), *resid = B<double>(ws.bresid, n), *slope = B<double>(ws.bslope, 4 * n);
    double *med = B<double>(ws.bmed, 4);
    void *medscr = ws.bmedscr.get<char>(median_scratch_bytes());
    uint8_t *low = B<uint8_t>(ws.blow, n), *kept = B<uint8_t>(ws.bkept, n), *within = B<uint8_t>(ws.bwithin, n);
    uint8_t *ok = B<uint8_t>(ws.bok, n);
    int *cold2 = B<int>(ws.bcold2, n), *colrow = B<int>(ws.bcolrow, n);
    Counters *cnt = B<Counters>(ws.bcnt, 1);
    float *conf = B<float>(ws.bconf, n);
    double *conf64 = B<double>(ws.bconf64, n);
    const double R = auto_edge_guard(c);
    const int freeze = c.prewarp_freeze_at;
    bool have_kept = false;
    step_mark("fine: outer iterations");
    for (int it = 0; it < c.outer_iters; it++) {
        if (it == 0) {
            ws.prof.mark(ST_SMOOTH, s);
            blur(ux, bt, sx, w, h, ws.taps_s_host.data(), ws.taps_s_r, s);
            blur(uy, bt, sy, w, h, ws.taps_s_host.data(), ws.taps_s_r, s);
        } else {
            // trusted pre-warp (phaseflow.py:326-355; frozen low mask: prewarp.rs:48-115)
            ws.prof.mark(ST_PREWARP, s);
            const uint8_t *lowm = low;
            if (freeze >= 1 && it > freeze && have_kept) {
                lowm = kept;
            } else {
                confidence(energy, resid, med, ux, uy, c.phase_tol, c.conf_min, 0.0, w, h, low, nullptr, nullptr, s);
                if (freeze >= 1 && it == freeze) {
                    dcopy(kept, low, n, s);
                    have_kept = true;
                }
            }
            int *colsum = B<int>(ws.bcolsum, column_summary_ints(w, h));
            column_nearest(lowm, w, h, cold2, colrow, colsum, s);
            within_distance(cold2, w, h, R, within, s);
            prewarp_trusted(lowm, within, w, h, ok, cnt, s);
            column_nearest(ok, w, h, cold2, colrow, colsum, s);
            prewarp_fill(ux, uy, slope, ok, cold2, colrow, cnt, w, h, fx, fy, s);
            ws.prof.mark(ST_SMOOTH, s);
            blur(fx, bt, sx, w, h, ws.taps_s_host.data(), ws.taps_s_r, s);
            blur(fy, bt, sy, w, h, ws.taps_s_host.data(), ws.taps_s_r, s);
        }
        ws.prof.mark(ST_WARP, s);
        warp(c2, w2, h2, shx, shy, sx, sy, warped, w, h, s);
        ws.prof.mark(ST_RESPONSES, s);
        fft_forward_padded(warped, w, h, spec, tmp, tmp2, s);
        fft_inverse_filtered(spec, hp, wp, bank, 0, F, MUL_G, true, tmp, tmp2, resp, s);
        ws.prof.mark(ST_SMOOTH, s);
        prefilter_real2(sx, sy, csx, csy, B<float>(ws.bpftmp, 2 * n), w, h, s);
        for (int in = 0; in < c.inner_iters; in++) {
            ws.prof.mark(ST_ACCUMULATE, s);
            accumulate(ux, uy, sx, sy, csx, csy, resp, refp, sums, w, h, F, s);
            ws.prof.mark(ST_MOMENTS, s);
            if (c.shape_order == 2)
                moments_solve2(sums, B<double>(ws.bmom2, moments2_scratch_doubles(n)), ux, uy, nux, nuy, energy, resid,
                               slope, w, h, (const double *)ws.btaps_w.p, ws.taps_w_r, s);
            else
            moments_solve(sums, ym, ux, uy, nux, nuy, energy, resid, slope, w, h, (const double *)ws.btaps_w.p,
                          ws.taps_w_host.data(), ws.taps_w_r, s);
            ws.prof.mark(ST_SOLVE, s);
            // the median itself is read only after an outer iteration's last inner one (the
            // confidence); before that the threshold's bounds decide the reset where they can.
            // OFX_EXACT_ENERGY_RESET set in the environment takes the exact median every time.
            if (in + 1 == c.inner_iters || std::getenv("OFX_EXACT_ENERGY_RESET")) {
                median(energy, n, medscr, med, s);
                energy_reset(energy, med, ux, uy, nux, nuy, resid, slope, n, s);
            } else {
                median_energy_reset(energy, n, medscr, med, ux, uy, nux, nuy, resid, slope, s);
            }
            std::swap(ux, nux);
            std::swap(uy, nuy);
        }
    }
    // the final field lives in ux / uy; keep the workspace's names pointing at it
    if (ux != (double *)ws.bux.p) {
        std::swap(ws.bux.p, ws.bnux.p);
        std::swap(ws.bux.bytes, ws.bnux.bytes);
        std::swap(ws.buy.p, ws.bnuy.p);
        std::swap(ws.buy.bytes, ws.bnuy.bytes);
    }
    const double Bw = edge_guard(c) > 0 ? 4.0 * c.sigma_w : 0.0;
    ws.prof.mark(ST_CONFIDENCE, s);
    confidence(energy, resid, med, ux, uy, c.phase_tol, c.conf_min, Bw, w, h, nullptr, conf, conf64, s);
}

// ---------------------------------------------------------------- the fine stage as a graph
// refine_device enqueues only kernels once its FFT plans, filter taps and buffers exist (no host
// wait, no copy by the runtime, no memset), so after one eager call with given sizes, settings,
// inputs and buffers, the second such call captures it into a HIP graph and every later one
// replays that graph: one launch instead of about a thousand. The key holds everything the
// captured launches depend on: the input pointers, every workspace buffer, the sizes and shift,
// the settings' bytes, the cached-pack state and the exact-reset switch. A call whose key differs
// runs eagerly (and warms its own key). Off with OFX_GRAPHS=0, with the stage profiler on, and when
// the inner iterations are odd in number (the result would end in the other field buffer).
bool graphs_enabled() {
    const char *v = std::getenv("OFX_GRAPHS");
    return !(v && v[0] == '0');
}

bool pack_cached(const Workspace &ws, int w, int h, const Bank &bank, bool cache_i1) {
    return cache_i1 && ws.pack_valid && ws.pack_w == w && ws.pack_h == h && same_bank(ws.pack_bank, bank);
}

std::vector<uint64_t> fine_key(const Workspace &ws, const float *i1, const float *i2, int w, int h, int w2, int h2,
                               long shx, long shy, const ofx_config &c, bool cache_i1) {
    std::vector<uint64_t> k = {(uint64_t)(uintptr_t)i1, (uint64_t)(uintptr_t)i2, (uint64_t)w, (uint64_t)h,
                               (uint64_t)w2, (uint64_t)h2, (uint64_t)shx, (uint64_t)shy, (uint64_t)cache_i1,
                               (uint64_t)pack_cached(ws, w, h, make_bank(c), cache_i1),
                               (uint64_t)(std::getenv("OFX_EXACT_ENERGY_RESET") != nullptr)};
    for (const DevBuf *b : {&ws.bspec, &ws.btmp, &ws.btmp2, &ws.bresp, &ws.brefp, &ws.bpack, &ws.bc2, &ws.bpftmp,
                            &ws.bux, &ws.buy, &ws.bnux, &ws.bnuy, &ws.bsx, &ws.bsy, &ws.bcsx, &ws.bcsy, &ws.bwarped,
                            &ws.bfx, &ws.bfy, &ws.bbt, &ws.bsums, &ws.bym, &ws.benergy, &ws.bresid, &ws.bslope,
                            &ws.bmed, &ws.bmedscr, &ws.blow, &ws.bkept, &ws.bwithin, &ws.bok, &ws.bcold2, &ws.bcolrow,
                            &ws.bcnt, &ws.bconf, &ws.bconf64, &ws.bcolsum, &ws.btaps_s, &ws.btaps_w, &ws.bmom2}) {
        k.push_back((uint64_t)(uintptr_t)b->p);
        k.push_back((uint64_t)b->bytes);
    }
    uint64_t cw[(sizeof(ofx_config) + 7) / 8] = {0};
    std::memcpy(cw, &c, sizeof(ofx_config));
    k.insert(k.end(), cw, cw + (sizeof(ofx_config) + 7) / 8);
    return k;
}

// refine_device, replayed from a graph where it can be (see above); results bit-identical.
void refine_graph(Workspace &ws, const float *i1, const float *i2, int w, int h, int w2, int h2, long shx, long shy,
                  const ofx_config &c, hipStream_t s, bool cache_i1) {
    const bool can = graphs_enabled() && ws.graphs_on && !ws.graphs_off && !ws.prof.on &&
                     (c.outer_iters * c.inner_iters) % 2 == 0;
    if (!can) {
        refine_device(ws, i1, i2, w, h, w2, h2, shx, shy, c, s, cache_i1);
        return;
    }
    const std::vector<uint64_t> key = fine_key(ws, i1, i2, w, h, w2, h2, shx, shy, c, cache_i1);
    for (auto &g : ws.graphs)
        if (g.exec && g.key == key) {
            step_mark("fine: replaying the captured fine stage");
            HIPCHECK(hipGraphLaunch(g.exec, s));
            ws.graph_launches++;
            // the cached reference pack a capture with cache_i1 computed is valid after a replay too
            if (cache_i1) {
                ws.pack_valid = true;
                ws.pack_w = w;
                ws.pack_h = h;
                ws.pack_bank = make_bank(c);
            }
            return;
        }
    bool warm = false;
    for (const auto &k : ws.warm_keys) warm = warm || k == key;
    if (!warm) {
        // first call with this key: eager; it creates the plans, taps and buffers a capture needs
        refine_device(ws, i1, i2, w, h, w2, h2, shx, shy, c, s, cache_i1);
        const std::vector<uint64_t> after = fine_key(ws, i1, i2, w, h, w2, h2, shx, shy, c, cache_i1);
        if (ws.warm_keys.size() >= 4) ws.warm_keys.erase(ws.warm_keys.begin());
        ws.warm_keys.push_back(after);
        return;
    }
    step_mark("fine: capturing the fine stage");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        refine_device(ws, i1, i2, w, h, w2, h2, shx, shy, c, s, cache_i1);
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        ws.graphs_off = true;
        throw;
    }
    HIPCHECK(hipStreamEndCapture(s, &graph));
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        ws.graphs_off = true;
        throw Error(OFX_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    auto &slot = ws.graphs[ws.graph_next];
    ws.graph_next ^= 1;
    if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
    slot.exec = exec;
    slot.key = key;
    ws.graph_captures++;
    HIPCHECK(hipGraphLaunch(exec, s));
    ws.graph_launches++;
}

// ---------------------------------------------------------------- tiled fine stage
// rust/crates/optflow-core/src/tile.rs:17-146 (geometry) and :225-500 (the tiled driver and
// join): cores of c.tile px partition the frame; each is refined on its extension (core plus
// tile_margin, clipped) against the window of image 2 its start field reaches, and only the core
// is kept. The confidence is formed afterwards from the joined energy (median over the frame) and
// the joined field (border rule), as for a whole frame.
struct TRect {
    int x0, y0, x1, y1;
    long long area() const { return (long long)(x1 - x0) * (y1 - y0); }
};

// per block: the finite range of (float) ux and uy over a rectangle of the frame
__global__ void __launch_bounds__(256) k_rect_minmax(const double *__restrict__ ux, const double *__restrict__ uy,
                                                     int w, int x0, int y0, int rw, int rh, double *__restrict__ part) {
    __shared__ double red[4][256];
    double lox = INFINITY, hix = -INFINITY, loy = INFINITY, hiy = -INFINITY;
    const long long total = (long long)rw * rh;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const int iy = (int)(k / rw), ix = (int)(k - (long long)iy * rw);
        const size_t i = (size_t)(y0 + iy) * w + (x0 + ix);
        const double a = (double)(float)ux[i], b = (double)(float)uy[i];
        if (isfinite(a)) {
            lox = fmin(lox, a);
            hix = fmax(hix, a);
        }
        if (isfinite(b)) {
            loy = fmin(loy, b);
            hiy = fmax(hiy, b);
        }
    }
    red[0][threadIdx.x] = lox;
    red[1][threadIdx.x] = hix;
    red[2][threadIdx.x] = loy;
    red[3][threadIdx.x] = hiy;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + off]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + off]);
            red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][threadIdx.x + off]);
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) part[4 * blockIdx.x + q] = red[q][0];
}

// the core of a refined tile (ew wide, at ex0, ey0 in the frame) into the joined planes
__global__ void k_write_core(const double *__restrict__ tux, const double *__restrict__ tuy,
                             const double *__restrict__ ten, const double *__restrict__ tres, int ew, int ex0, int ey0,
                             int cx0, int cy0, int cw, int w, double *__restrict__ jx, double *__restrict__ jy,
                             double *__restrict__ je, double *__restrict__ jr) {
    const int dx = blockIdx.x * blockDim.x + threadIdx.x, y = cy0 + blockIdx.y;
    if (dx >= cw) return;
    const int x = cx0 + dx;
    const size_t l = (size_t)(y - ey0) * ew + (x - ex0), i = (size_t)y * w + x;
    jx[i] = tux[l];
    jy[i] = tuy[l];
    je[i] = ten[l];
    jr[i] = tres[l];
}

// Tile groups refined side by side in the tiled fine stage: by default the lane's stream and two
// more, which with the stream of the caller keeps to the four hardware queues a process gets by
// default (ofx_set_schedule changes it, 1 .. MAX_TILE_GROUPS).
constexpr int MAX_TILE_GROUPS = 8;

// The fine stage of a pair whose dense start is in ws.bux / ws.buy (w x h): the whole frame, or
// tiled when the frame has a side over c.tile (track.rs:580-598). Leaves the field in ws.bux /
// ws.buy, energy and residual in ws.benergy / ws.bresid, the confidence in ws.bconf / ws.bconf64.
void fine_stage(Workspace &ws, const float *i1, const float *i2, int w, int h, const ofx_config &c, hipStream_t s,
                bool cache_i1) {
    if (c.tile <= 0 || std::max(w, h) <= c.tile) {
        refine_graph(ws, i1, i2, w, h, w, h, 0, 0, c, s, cache_i1);
        return;
    }
    const size_t n = (size_t)w * h;
    double *fsx = B<double>(ws.bfsx, n), *fsy = B<double>(ws.bfsy, n);
    dcopy(fsx, ws.bux.p, n * sizeof(double), s);
    dcopy(fsy, ws.buy.p, n * sizeof(double), s);
    double *jx = B<double>(ws.bjx, n), *jy = B<double>(ws.bjy, n), *je = B<double>(ws.bje, n),
           *jr = B<double>(ws.bjr, n);
    // every plane the tiles use and the joined result is copied back into holds the whole frame
    // before the first tile: refine_device swaps bux / bnux, and growing a buffer later, while the
    // last tile's kernels still read it on the stream, would free memory in use
    for (DevBuf *b : {&ws.bux, &ws.buy, &ws.bnux, &ws.bnuy, &ws.benergy, &ws.bresid}) (void)B<double>(*b, n);
    const int t = c.tile, m = tile_margin(c);
    // tile.rs:38-76: row by row, then the largest extension first (stable)
    std::vector<std::pair<TRect, TRect>> ts;  // (core, ext)
    auto axis = [](int nn, int tt, int mm, int k, int &c0, int &c1, int &e0, int &e1) {
        c0 = k * tt;
        c1 = std::min(nn, c0 + tt);
        e0 = std::max(0, c0 - mm);
        e1 = std::min(nn, c1 + mm);
    };
    for (int ky = 0; ky < (h + t - 1) / t; ky++)
        for (int kx = 0; kx < (w + t - 1) / t; kx++) {
            TRect core, ext;
            axis(w, t, m, kx, core.x0, core.x1, ext.x0, ext.x1);
            axis(h, t, m, ky, core.y0, core.y1, ext.y0, ext.y1);
            ts.push_back({core, ext});
        }
    std::vector<size_t> order(ts.size());
    for (size_t k = 0; k < ts.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return ts[a].second.area() > ts[b].second.area(); });
    // tile.rs:112-146: each extension moved by the finite range of its (f32) start field; the
    // ranges of every tile come from the start field alone, so they are found first, in one pass
    const int nblk = 256, nt = (int)ts.size();
    double *part = B<double>(ws.bpart, (size_t)4 * nblk * nt);
    for (int k = 0; k < nt; k++) {
        const TRect ext = ts[k].second;
        hipLaunchKernelGGL(k_rect_minmax, dim3(nblk), dim3(256), 0, s, (const double *)fsx, (const double *)fsy, w,
                           ext.x0, ext.y0, ext.x1 - ext.x0, ext.y1 - ext.y0, part + (size_t)4 * nblk * k);
        LAUNCHCHECK();
    }
    std::vector<double> hpart((size_t)4 * nblk * nt);
    copy_d2h(ws.pin, hpart.data(), part, hpart.size() * sizeof(double), s);
    std::vector<TRect> wins(nt);
    for (int k = 0; k < nt; k++) {
        const TRect ext = ts[k].second;
        const double *hp = hpart.data() + (size_t)4 * nblk * k;
        double lox = INFINITY, hix = -INFINITY, loy = INFINITY, hiy = -INFINITY;
        for (int b = 0; b < nblk; b++) {
            lox = std::min(lox, hp[4 * b]);
            hix = std::max(hix, hp[4 * b + 1]);
            loy = std::min(loy, hp[4 * b + 2]);
            hiy = std::max(hiy, hp[4 * b + 3]);
        }
        if (lox > hix) lox = hix = 0.0;
        if (loy > hiy) loy = hiy = 0.0;
        auto span = [&](int first, int end, double lo, double hi, int nn, int &a, int &b) {
            double fa = std::floor(first + lo) - m, fb = std::ceil((end - 1) + hi) + 1.0 + m;
            fa = std::min(std::max(fa, 0.0), (double)(nn - 1));
            a = (int)fa;
            fb = std::min(std::max(fb, (double)(a + 1)), (double)nn);
            b = (int)fb;
        };
        span(ext.x0, ext.x1, lox, hix, w, wins[k].x0, wins[k].x1);
        span(ext.y0, ext.y1, loy, hiy, h, wins[k].y0, wins[k].y1);
        const int ww = wins[k].x1 - wins[k].x0, wh = wins[k].y1 - wins[k].y0;
        if (ww > MAX_I2_SIDE || wh > MAX_I2_SIDE)
            throw Error(OFX_ETOOLARGE, "the image 2 window of a tile, " + std::to_string(ww) + "x" + std::to_string(wh) +
                                           ", exceeds the " + std::to_string(MAX_I2_SIDE) +
                                           " px side the spline rows hold (the start field spans too far)");
    }
    // The tiles are independent refines (tile.rs:225-330): they run in ws.tile_groups groups, the
    // group g taking the tiles order[g], order[g + G], ... (largest first), group 0 on the lane's
    // stream and workspace, the others on streams and workspaces of their own that fork from the
    // lane's stream here and join it before the joined planes are read. Each group's workspace
    // is used by one stream only; the cores the tiles write into jx .. jr are disjoint.
    const int G = std::max(1, std::min(ws.tile_groups, nt));
    while ((int)ws.sub.size() < G - 1) {
        ws.sub.emplace_back(new Workspace);
        hipStream_t t;
        HIPCHECK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        ws.sub_streams.push_back(t);
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ws.sub_done.push_back(e);
    }
    if (!ws.fork) HIPCHECK(hipEventCreateWithFlags(&ws.fork, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(ws.fork, s));
    for (int g = 1; g < G; g++) {
        HIPCHECK(hipStreamWaitEvent(ws.sub_streams[g - 1], ws.fork, 0));
        ws.sub[g - 1]->prof.on = ws.prof.on;
    }
    for (int q = 0; q < nt; q++) {
        step_mark("fine: a tile");
        const int g = q % G;
        Workspace &tw = g == 0 ? ws : *ws.sub[g - 1];
        hipStream_t ts_ = g == 0 ? s : ws.sub_streams[g - 1];
        const size_t k = order[q];
        const TRect core = ts[k].first, ext = ts[k].second, win = wins[k];
        const int ew = ext.x1 - ext.x0, eh = ext.y1 - ext.y0;
        const size_t ne = (size_t)ew * eh;
        const int ww = win.x1 - win.x0, wh = win.y1 - win.y0;
        float *t1 = B<float>(tw.btile1, ne), *t2 = B<float>(tw.btile2, (size_t)ww * wh);
        dcopy2d(t1, ew * sizeof(float), i1 + (size_t)ext.y0 * w + ext.x0, w * sizeof(float), ew * sizeof(float), eh,
                ts_);
        dcopy2d(t2, ww * sizeof(float), i2 + (size_t)win.y0 * w + win.x0, w * sizeof(float), ww * sizeof(float), wh,
                ts_);
        double *sx = B<double>(tw.bux, ne), *sy = B<double>(tw.buy, ne);
        dcopy2d(sx, ew * sizeof(double), fsx + (size_t)ext.y0 * w + ext.x0, w * sizeof(double),
                ew * sizeof(double), eh, ts_);
        dcopy2d(sy, ew * sizeof(double), fsy + (size_t)ext.y0 * w + ext.x0, w * sizeof(double),
                ew * sizeof(double), eh, ts_);
        refine_device(tw, t1, t2, ew, eh, ww, wh, ext.x0 - win.x0, ext.y0 - win.y0, c, ts_, false);
        const int cw = core.x1 - core.x0;
        hipLaunchKernelGGL(k_write_core, dim3((cw + 255) / 256, core.y1 - core.y0), dim3(256), 0, ts_,
                           (const double *)tw.bux.p, (const double *)tw.buy.p, (const double *)tw.benergy.p,
                           (const double *)tw.bresid.p, ew, ext.x0, ext.y0, core.x0, core.y0, cw, w, jx, jy, je, jr);
        LAUNCHCHECK();
    }
    for (int g = 1; g < G; g++) {
        HIPCHECK(hipEventRecord(ws.sub_done[g - 1], ws.sub_streams[g - 1]));
        HIPCHECK(hipStreamWaitEvent(s, ws.sub_done[g - 1], 0));
    }
    if (ws.prof.on) {
        // the stage times of the other groups, measured on their streams, join the lane's
        for (int g = 1; g < G; g++) {
            Profiler &sp = ws.sub[g - 1]->prof;
            sp.flush();
            for (int i = 0; i < ST_COUNT; i++) {
                ws.prof.ms[i] += sp.ms[i];
                ws.prof.calls[i] += sp.calls[i];
            }
            sp.reset();
        }
    }
    double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
    double *energy = B<double>(ws.benergy, n), *resid = B<double>(ws.bresid, n);
    dcopy(ux, jx, n * sizeof(double), s);
    dcopy(uy, jy, n * sizeof(double), s);
    dcopy(energy, je, n * sizeof(double), s);
    dcopy(resid, jr, n * sizeof(double), s);
    double *med = B<double>(ws.bmed, 4);
    ws.prof.mark(ST_CONFIDENCE, s);
    median(energy, n, ws.bmedscr.get<char>(median_scratch_bytes()), med, s);
    const double Bw = edge_guard(c) > 0 ? 4.0 * c.sigma_w : 0.0;
    confidence(energy, resid, med, ux, uy, c.phase_tol, c.conf_min, Bw, w, h, nullptr, B<float>(ws.bconf, n),
               B<double>(ws.bconf64, n), s);
}

// Validity, guard and strain (track.py:56-108) of the field (ux, uy): low = conf < conf_min, or the
// given low mask (the hybrid sequence's, hybrid.rs:495-503).
void finish_device(Workspace &ws, const double *ux, const double *uy, const float *conf, const uint8_t *low_in,
                   int w, int h, const uint8_t *mask, const ofx_config &c, ofx_result *o, hipStream_t s) {
    const size_t n = (size_t)w * h;
    uint8_t *low = B<uint8_t>(ws.blow, n), *near = B<uint8_t>(ws.bnear, n);
    const double guard = edge_guard(c);
    const int guard_on = guard > 0;
    ws.prof.mark(ST_GUARD, s);
    if (low_in) dcopy(low, low_in, n, s);
    else low_mask(conf, c.conf_min, n, low, s);
    if (guard_on) {
        uint8_t *large = B<uint8_t>(ws.blarge, n), *src = B<uint8_t>(ws.bsrc, n);
        int *parent = B<int>(ws.bparent, n);
        unsigned int *sizes = B<unsigned int>(ws.bsizes, n);
        int *cold2 = B<int>(ws.bcold2, n), *colrow = B<int>(ws.bcolrow, n);
        large_components(low, w, h, edge_guard_min_area(c), parent, sizes, large, s);
        guard_sources(large, mask, w, h, src, s);
        column_nearest(src, w, h, cold2, colrow, B<int>(ws.bcolsum, column_summary_ints(w, h)), s);
        within_distance(cold2, w, h, guard, near, s);
    }
    double *mx = B<double>(ws.bmx, n), *my = B<double>(ws.bmy, n), *yt = B<double>(ws.bytmp, 11 * n);
    validity(low, mask, near, guard, guard_on, ux, uy, w, h, mx, my, s);
    ws.prof.mark(ST_STRAIN, s);
    strain(mx, my, w, h, c.strain_win, yt, o->ux, o->uy, o->valid, o->exx, o->eyy, o->exy, o->gmax, o->vol, s);
}

__global__ void k_fill(double *a, double v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_u8(uint8_t *a, uint8_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_i32(int *a, int v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_to_f32(const double *a, float *o, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)a[i];
}
__global__ void k_to_f64(const float *a, double *o, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (double)a[i];
}
void to_f32(const double *a, float *o, size_t n, hipStream_t s) {
    if (!o) return;
    hipLaunchKernelGGL(k_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, a, o, n);
    LAUNCHCHECK();
}
void to_f64(const float *a, double *o, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f64, dim3((n + 255) / 256), dim3(256), 0, s, a, o, n);
    LAUNCHCHECK();
}

__global__ void k_min_inplace(double *a, const double *b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = fmin(a[i], b[i]);
}
__global__ void k_min_pair(const double *a, const double *b, double *o, float *of, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        o[i] = fmin(a[i], b[i]);
        of[i] = (float)o[i];
    }
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// The coarse stage of a pair: the POC pyramid (poc.rs:694-766), or with coarse_learned (and no dense
// guess) the level-0 grid started from the learned proposal, the pyramid running instead when
// fewer than coarse_gate of the grid points confirm the proposal (a POC peak of at least
// poc_peak_min and a result within poc_win / 8 px of it). The outcome goes to o->coarse_*.
void coarse_stage(Workspace &ws, const float *i1, const float *i2, int w, int h, const ofx_config &c,
                  const double *init_x, const double *init_y, double *ux, double *uy, ofx_result *o, hipStream_t s) {
    o->coarse_confirmed = -1.0;
    o->coarse_fallback = 0;
    if (!c.coarse_learned || init_x) {
        coarse_dense(i1, i2, w, h, c, init_x, init_y, ux, uy, ws.coarse, s);
        return;
    }
    if (!ws.net) throw Error(OFX_EINVAL, "coarse_learned needs a network: load one with ofx_coarse_net_load");
    if (!ws.nacts) ws.nacts.reset(new NetActs);
    NetActs &A = *ws.nacts;
    net_propose(*ws.net, A, i1, i2, w, h, s);
    int confirmed = 0;
    const int np = coarse_dense_from(
        i1, i2, w, h, c,
        [&](int x0, int y0, int step, int nx, int ny, double *ix, double *iy) {
            prop_to_grid(A, x0, y0, step, nx, ny, ix, iy, s);
        },
        c.poc_win / 8.0, ux, uy, ws.coarse, &confirmed, s);
    o->coarse_confirmed = (double)confirmed / np;
    if (o->coarse_confirmed < c.coarse_gate) {
        o->coarse_fallback = 1;
        coarse_dense(i1, i2, w, h, c, nullptr, nullptr, ux, uy, ws.coarse, s);
    }
}

// The pair pipeline (track.py:115-146) on device inputs; init as double (or NULL); outputs to
// device result arrays. Timing by events on the lane's stream.
void pair_device(Workspace &ws, const float *i1, const float *i2, int w, int h, const uint8_t *mask,
                 const double *init_x, const double *init_y, const ofx_config &c, ofx_result *o, hipStream_t s) {
    const size_t n = (size_t)w * h;
    for (auto &e : ws.ev)
        if (!e) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipEventRecord(ws.ev[0], s));
    ws.prof.mark(ST_COARSE, s);
    double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
    step_mark("pair: coarse stage");
    coarse_stage(ws, i1, i2, w, h, c, init_x, init_y, ux, uy, o, s);
    HIPCHECK(hipEventRecord(ws.ev[1], s));
    step_mark("pair: fine stage");
    fine_stage(ws, i1, i2, w, h, c, s, false);
    step_mark("pair: guard and strain");
    HIPCHECK(hipEventRecord(ws.ev[2], s));
    ux = (double *)ws.bux.p;
    uy = (double *)ws.buy.p;
    finish_device(ws, ux, uy, (const float *)ws.bconf.p, nullptr, w, h, mask, c, o, s);
    to_f32(ux, o->raw_ux, n, s);
    to_f32(uy, o->raw_uy, n, s);
    if (o->conf) dcopy(o->conf, ws.bconf.p, n * sizeof(float), s);
    ws.prof.mark(ST_END, s);
    HIPCHECK(hipEventRecord(ws.ev[3], s));
}

void record_times(Workspace &ws, ofx_result *o) {
    step_mark("pair: waiting for the end event");
    wait_event(ws.ev[3]);
    ws.prof.flush();
    o->t_coarse = elapsed(ws.ev[0], ws.ev[1]) * 1e-3;
    o->t_fine = elapsed(ws.ev[1], ws.ev[2]) * 1e-3;
    o->t_strain = elapsed(ws.ev[2], ws.ev[3]) * 1e-3;
}

// Device staging for the host-pointer API: 10 float planes + the valid plane.
struct Staged {
    ofx_result d{};
};

Staged stage_outputs(Workspace &ws, size_t n, const ofx_result *host) {
    Staged st;
    float *base = (float *)ws.bout.get<char>(10 * n * sizeof(float) + n);
    float **dst[10] = {&st.d.ux, &st.d.uy, &st.d.conf, &st.d.raw_ux, &st.d.raw_uy,
                       &st.d.exx, &st.d.eyy, &st.d.exy, &st.d.gmax, &st.d.vol};
    float *const src[10] = {host->ux, host->uy, host->conf, host->raw_ux, host->raw_uy,
                            host->exx, host->eyy, host->exy, host->gmax, host->vol};
    for (int k = 0; k < 10; k++) *dst[k] = src[k] ? base + k * n : nullptr;
    st.d.valid = host->valid ? (uint8_t *)(base + 10 * n) : nullptr;
    return st;
}

void download(Workspace &ws, const Staged &st, ofx_result *host, size_t n, hipStream_t s) {
    float *const d[10] = {st.d.ux, st.d.uy, st.d.conf, st.d.raw_ux, st.d.raw_uy,
                          st.d.exx, st.d.eyy, st.d.exy, st.d.gmax, st.d.vol};
    float *const hst[10] = {host->ux, host->uy, host->conf, host->raw_ux, host->raw_uy,
                            host->exx, host->eyy, host->exy, host->gmax, host->vol};
    for (int k = 0; k < 10; k++)
        if (hst[k]) copy_d2h(ws.pin, hst[k], d[k], n * sizeof(float), s);
    step_mark("download: the result copies");
    if (host->valid) copy_d2h(ws.pin, host->valid, st.d.valid, n, s);
}

// the argument checks of track_pair (track.rs:655-663, track.py:115-125)
void check_pair_args(const float *i1, const float *i2, int w, int h, const uint8_t *mask, const float *init_ux,
                     const float *init_uy, const ofx_config &c, const ofx_result *out) {
    if (!i1 || !i2 || !out) throw Error(OFX_EINVAL, "input or output pointer is NULL");
    if ((init_ux == nullptr) != (init_uy == nullptr)) throw Error(OFX_EINVAL, "init needs both components");
    check_window_impl(w, h, c);
    if (std::min(w, h) < c.poc_win) throw Error(OFX_EINVAL, "image smaller than poc_win (" + std::to_string(c.poc_win) + ")");
    if (mask) {
        const size_t n = (size_t)w * h;
        bool any = false;
        for (size_t i = 0; i < n && !any; i++) any = mask[i] != 0;
        if (!any) throw Error(OFX_EINVAL, "ROI mask is empty");
    }
}

// One pair from host inputs to host outputs on a lane the caller holds.
void track_pair_on_lane(Lane &lane, const float *i1, const float *i2, int w, int h, const uint8_t *mask,
                        const float *init_ux, const float *init_uy, const ofx_config &c, ofx_result *out) {
    Workspace &ws = *lane.ws;
    hipStream_t s = lane.stream;
    const size_t n = (size_t)w * h;
    float *d1 = B<float>(ws.bi1, n), *d2 = B<float>(ws.bi2, n);
    copy_h2d(ws.pin, d1, i1, n * sizeof(float), s);
    copy_h2d(ws.pin, d2, i2, n * sizeof(float), s);
    uint8_t *dm = nullptr;
    if (mask) {
        dm = B<uint8_t>(ws.bmask, n);
        copy_h2d(ws.pin, dm, mask, n, s);
    }
    double *ix = nullptr, *iy = nullptr;
    if (init_ux) {
        float *fx = B<float>(ws.binx, n), *fy = B<float>(ws.biny, n);
        copy_h2d(ws.pin, fx, init_ux, n * sizeof(float), s);
        copy_h2d(ws.pin, fy, init_uy, n * sizeof(float), s);
        ix = B<double>(ws.bstart_x, n);
        iy = B<double>(ws.bstart_y, n);
        to_f64(fx, ix, n, s);
        to_f64(fy, iy, n, s);
    }
    Staged st = stage_outputs(ws, n, out);
    pair_device(ws, d1, d2, w, h, dm, ix, iy, c, &st.d, s);
    record_times(ws, &st.d);
    download(ws, st, out, n, s);
    out->t_coarse = st.d.t_coarse;
    out->t_fine = st.d.t_fine;
    out->t_strain = st.d.t_strain;
    out->coarse_confirmed = st.d.coarse_confirmed;
    out->coarse_fallback = st.d.coarse_fallback;
    out->reference = 0;
    out->mean_conf = 0;
}

template <class F>
int guarded(F &&f) {
    // the outermost call on this thread times itself for the watchdog
    static thread_local int depth = 0;
    struct Timed {
        bool outer;
        Timed() : outer(depth++ == 0) {
            if (outer) {
                Progress &p = progress();
                p.call_id.fetch_add(1);
                p.pending.store(-1);
                p.call_since_ms.store(now_ms());
            }
        }
        ~Timed() {
            depth--;
            if (outer) progress().call_since_ms.store(0);
        }
    } timed;
    try {
        f();
        return OFX_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return OFX_EDEVICE;
    }
}


}  // namespace

int guarded_call(const std::function<void()> &f) { return guarded(f); }
void validate_config(const ofx_config *cfg) { validate(cfg); }
double guard_band_px(const ofx_config &c) { return edge_guard(c); }

namespace {

// A sequence on one lane (rust/crates/optflow-core/src/track.rs:1112-1225 Sequence::new and
// step; track.py:260-315 for seq_mode "chain", track/hybrid.rs:262-540 and track.py:383-505 for
// "hybrid"). Its state lives in the lane's workspace: the reference frame, the previous frame,
// u_0r (frame 0 -> reference), the frame-0 link confidence c0r, the previous leg's raw field and
// confidence, and for hybrid the per-pixel anchor, carried field, fallback count and adjacent-leg
// start. The lane stays held from begin to the last step, so nothing else touches that state.
struct Seq {
    ofx_engine *eng;
    Lane *lane;
    Workspace &ws;
    hipStream_t s;
    ofx_config c;
    int w, h;
    size_t n;
    uint8_t *dm = nullptr;
    double Bw;
    int k = 0;  // index of the last frame seen
    int ref_idx = 0;
    bool have_prev = false, have_prev_adj = false;

    Seq(ofx_engine *e, Lane *l, const ofx_config &cfg, int w_, int h_)
        : eng(e), lane(l), ws(*l->ws), s(l->stream), c(cfg), w(w_), h(h_), n((size_t)w_ * h_),
          Bw(edge_guard(cfg) > 0 ? 4.0 * cfg.sigma_w : 0.0) {}

    // one leg a -> b from the start (ix, iy) or the coarse stage alone; its raw field and
    // confidence go to (ox, oy, oc); returns the coarse and fine seconds
    // (cache_a: a is the reference frame, whose pack the workspace may keep between steps)
    std::pair<double, double> run_leg_to(const float *a, const float *b, const double *ix, const double *iy, double *ox,
                                         double *oy, double *oc, bool cache_a = false) {
        for (auto &ev : ws.ev)
            if (!ev) HIPCHECK(hipEventCreate(&ev));
        HIPCHECK(hipEventRecord(ws.ev[0], s));
        ws.prof.mark(ST_COARSE, s);
        double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
        step_mark("sequence: coarse stage of a leg");
        coarse_dense(a, b, w, h, c, ix, iy, ux, uy, ws.coarse, s);
        step_mark("sequence: fine stage of a leg");
        HIPCHECK(hipEventRecord(ws.ev[1], s));
        fine_stage(ws, a, b, w, h, c, s, cache_a);
        HIPCHECK(hipEventRecord(ws.ev[2], s));
        dcopy(ox, ws.bux.p, n * sizeof(double), s);
        dcopy(oy, ws.buy.p, n * sizeof(double), s);
        dcopy(oc, ws.bconf64.p, n * sizeof(double), s);
        step_mark("sequence: waiting for a leg");
        wait_event(ws.ev[2]);
        return std::make_pair(elapsed(ws.ev[0], ws.ev[1]) * 1e-3, elapsed(ws.ev[1], ws.ev[2]) * 1e-3);
    }
    // the direct leg reference -> current into (lx, ly, lc), from the previous leg's raw field
    std::pair<double, double> run_leg(bool with_init) {
        return run_leg_to((const float *)ws.bframe_ref.p, (const float *)ws.bi2.p,
                          with_init ? (const double *)ws.bprev_x.p : nullptr,
                          with_init ? (const double *)ws.bprev_y.p : nullptr, (double *)ws.bstart_x.p,
                          (double *)ws.bstart_y.p, (double *)ws.bleg_c.p, true);
    }
    // mean link confidence of the direct leg over the reportable ROI pixels (track.rs:986-1023)
    double leg_mean() {
        double *macc = B<double>(ws.bmean, 2 + 2 * (size_t)h + n);
        mean_conf((const double *)ws.bleg_c.p, (const double *)ws.bstart_x.p, (const double *)ws.bstart_y.p,
                  (const double *)ws.bu0x.p, (const double *)ws.bu0y.p, dm, Bw, w, h, macc, s);
        double acc[2];
        step_mark("sequence: the mean confidence");
        copy_d2h(ws.pin, acc, macc, 2 * sizeof(double), s);
        return acc[1] > 0 ? acc[0] / acc[1] : 0.0;
    }

    void begin(const float *frame0, bool device_frame, const uint8_t *mask) {
        if (mask) {
            dm = B<uint8_t>(ws.bmask, n);
            copy_h2d(ws.pin, dm, mask, n, s);
        }
        float *fref = B<float>(ws.bframe_ref, n), *fprev = B<float>(ws.bframe_prev, n);
        (void)B<float>(ws.bi2, n);
        ws.pack_valid = false;  // a new reference frame
        if (device_frame)
            dcopy(fref, frame0, n * sizeof(float), s);
        else
            copy_h2d(ws.pin, fref, frame0, n * sizeof(float), s);
        dcopy(fprev, fref, n * sizeof(float), s);
        double *u0x = B<double>(ws.bu0x, n), *u0y = B<double>(ws.bu0y, n), *c0r = B<double>(ws.bc0r, n);
        for (DevBuf *b : {&ws.bcx, &ws.bcy, &ws.bconf0, &ws.bca, &ws.bprev_x, &ws.bprev_y, &ws.bprev_c, &ws.bleg_c,
                          &ws.bstart_x, &ws.bstart_y})
            (void)B<double>(*b, n);
        hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, u0x, 0.0, n);
        hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, u0y, 0.0, n);
        hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, c0r, 1.0, n);
        LAUNCHCHECK();
        if (c.seq_hybrid) {
            uint8_t *anch = B<uint8_t>(ws.banch, n), *good_prev = B<uint8_t>(ws.bgood_prev, n);
            double *upx = B<double>(ws.bupx, n), *upy = B<double>(ws.bupy, n), *caa = B<double>(ws.bca_adj, n);
            int *nfb = B<int>(ws.bnfb, n);
            hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, anch, (uint8_t)1, n);
            hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, good_prev, (uint8_t)1, n);
            hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, nfb, 0, n);
            hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, upx, 0.0, n);
            hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, upy, 0.0, n);
            hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, caa, 0.0, n);
            LAUNCHCHECK();
        }
        wait_stream(s);
        k = 0;
        ref_idx = 0;
        have_prev = have_prev_adj = false;
    }

    void step(const float *frame, bool device_frame, ofx_result *out, bool device_out) {
        k++;
        if (device_frame)
            dcopy(ws.bi2.p, frame, n * sizeof(float), s);
        else
            copy_h2d(ws.pin, ws.bi2.p, frame, n * sizeof(float), s);
        if (c.seq_hybrid)
            step_hybrid(out, device_out);
        else
            step_chain(out, device_out);
    }

    // the result planes: staged in the workspace and downloaded for a host result, written in
    // place for a device one
    Staged outputs(ofx_result *out, bool device_out) {
        if (device_out) {
            Staged st;
            st.d = *out;
            return st;
        }
        return stage_outputs(ws, n, out);
    }
    void deliver(const Staged &st, ofx_result *out, bool device_out) {
        if (device_out)
            wait_stream(s);
        else
            download(ws, st, out, n, s);
    }

    void step_chain(ofx_result *out, bool device_out) {
        float *fref = (float *)ws.bframe_ref.p, *fprev = (float *)ws.bframe_prev.p, *fcur = (float *)ws.bi2.p;
        double *u0x = (double *)ws.bu0x.p, *u0y = (double *)ws.bu0y.p, *c0r = (double *)ws.bc0r.p;
        double *cx = (double *)ws.bcx.p, *cy = (double *)ws.bcy.p, *conf0 = (double *)ws.bconf0.p;
        double *ca = (double *)ws.bca.p, *lc = (double *)ws.bleg_c.p;
        double *px = (double *)ws.bprev_x.p, *py = (double *)ws.bprev_y.p, *pc = (double *)ws.bprev_c.p;
        double *lx = (double *)ws.bstart_x.p, *ly = (double *)ws.bstart_y.p;
        auto tl = run_leg(have_prev);
        double cm = leg_mean();
        if (cm < c.ref_update_conf && ref_idx != k - 1) {
            // prev is the leg old reference -> k - 1: its confidence extends frame 0's link
            conf_at(pc, u0x, u0y, w, h, ca, s);
            hipLaunchKernelGGL(k_min_inplace, dim3((n + 255) / 256), dim3(256), 0, s, c0r, (const double *)ca, n);
            compose(u0x, u0y, px, py, w, h, cx, cy, s);
            dcopy(u0x, cx, n * sizeof(double), s);
            dcopy(u0y, cy, n * sizeof(double), s);
            dcopy(fref, fprev, n * sizeof(float), s);
            ws.pack_valid = false;  // the reference moved
            ref_idx = k - 1;
            tl = run_leg(false);
            cm = leg_mean();
        }
        step_mark("chain: compose and finish");
        compose(u0x, u0y, lx, ly, w, h, cx, cy, s);
        conf_at(lc, u0x, u0y, w, h, ca, s);
        hipLaunchKernelGGL(k_min_pair, dim3((n + 255) / 256), dim3(256), 0, s, (const double *)c0r, (const double *)ca,
                           conf0, (float *)ws.bconf.p, n);
        LAUNCHCHECK();
        Staged st = outputs(out, device_out);
        finish_device(ws, cx, cy, (const float *)ws.bconf.p, nullptr, w, h, dm, c, &st.d, s);
        to_f32(cx, st.d.raw_ux, n, s);
        to_f32(cy, st.d.raw_uy, n, s);
        if (st.d.conf) dcopy(st.d.conf, ws.bconf.p, n * sizeof(float), s);
        ws.prof.mark(ST_END, s);
        HIPCHECK(hipEventRecord(ws.ev[3], s));
        deliver(st, out, device_out);
        ws.prof.flush();
        wait_event(ws.ev[3]);
        out->t_coarse = tl.first;
        out->t_fine = tl.second;
        out->t_strain = elapsed(ws.ev[2], ws.ev[3]) * 1e-3;
        out->reference = ref_idx;
        out->mean_conf = cm;
        // this leg becomes prev
        dcopy(px, lx, n * sizeof(double), s);
        dcopy(py, ly, n * sizeof(double), s);
        dcopy(pc, lc, n * sizeof(double), s);
        dcopy(fprev, fcur, n * sizeof(float), s);
        have_prev = true;
        wait_stream(s);
    }

    // seq_mode "hybrid": the direct leg of the chain plus, per pixel, the gated adjacent-frame
    // fallback. The host reads back the few counters that decide which legs and gates run.
    void step_hybrid(ofx_result *out, bool device_out) {
        float *fref = (float *)ws.bframe_ref.p, *fprev = (float *)ws.bframe_prev.p, *fcur = (float *)ws.bi2.p;
        double *u0x = (double *)ws.bu0x.p, *u0y = (double *)ws.bu0y.p, *c0r = (double *)ws.bc0r.p;
        double *cx = (double *)ws.bcx.p, *cy = (double *)ws.bcy.p, *ca = (double *)ws.bca.p;
        double *px = (double *)ws.bprev_x.p, *py = (double *)ws.bprev_y.p, *pc = (double *)ws.bprev_c.p;
        double *lc = (double *)ws.bleg_c.p, *lx = (double *)ws.bstart_x.p, *ly = (double *)ws.bstart_y.p;
        double *legc = B<double>(ws.bconf0, n);
        uint8_t *anch = B<uint8_t>(ws.banch, n), *good_prev = B<uint8_t>(ws.bgood_prev, n);
        double *upx = B<double>(ws.bupx, n), *upy = B<double>(ws.bupy, n);
        int *nfb = B<int>(ws.bnfb, n);
        double *pajx = B<double>(ws.bpadj_x, n), *pajy = B<double>(ws.bpadj_y, n);
        double *ajx = B<double>(ws.badj_x, n), *ajy = B<double>(ws.badj_y, n), *ajc = B<double>(ws.badj_c, n);
        double *link = B<double>(ws.blink, n);
        uint8_t *flags = B<uint8_t>(ws.bflags, n);
        double *ax = B<double>(ws.bax, n), *ay = B<double>(ws.bay, n);
        double *cax = B<double>(ws.bcax, n), *cay = B<double>(ws.bcay, n), *caa = B<double>(ws.bca_adj, n);
        uint8_t *gate0 = B<uint8_t>(ws.bgate0, n), *gate = B<uint8_t>(ws.bgate, n), *hlow = B<uint8_t>(ws.bhlow, n);
        int *cnt = B<int>(ws.bhcnt, HYB_NCOUNTERS);
        void *gscr = ws.bgscr.get<char>(gate_scratch_bytes(w, h));
        double *hux = B<double>(ws.bhux, n), *huy = B<double>(ws.bhuy, n);
        int32_t *dfb = B<int32_t>(ws.bfb, n);
        auto counters = [&]() {
            std::vector<int> hc(HYB_NCOUNTERS);
            step_mark("hybrid: the counters");
            copy_d2h(ws.pin, hc.data(), cnt, HYB_NCOUNTERS * sizeof(int), s);
            return hc;
        };
        const bool gate_on = c.fb_min_support > 0;
        const double radius = fb_radius_px(c);
        auto tl = run_leg(have_prev);
        double cm = leg_mean();
        if (cm < c.ref_update_conf && ref_idx != k - 1) {
            conf_at(pc, u0x, u0y, w, h, ca, s);
            hyb_refmove(good_prev, ca, c0r, anch, n, s);
            dcopy(u0x, upx, n * sizeof(double), s);
            dcopy(u0y, upy, n * sizeof(double), s);
            dcopy(fref, fprev, n * sizeof(float), s);
            ws.pack_valid = false;  // the reference moved
            ref_idx = k - 1;
            tl = run_leg(false);
            cm = leg_mean();
        }
        compose(u0x, u0y, lx, ly, w, h, cx, cy, s);
        conf_at(lc, u0x, u0y, w, h, legc, s);
        hyb_zero_counters(cnt, s);
        hyb_classify(cx, cy, legc, c0r, anch, dm, upx, upy, nfb, c.conf_min, c.fb_max, ref_idx + 1 < k, link, flags, cnt,
                     n, s);
        bool have_adj = false;
        if (counters()[HYB_ANY_NEED]) {
            const double *rx = lx, *ry = ly, *rc = lc;
            if (ref_idx != k - 1) {
                auto ta = run_leg_to(fprev, fcur, have_prev_adj ? pajx : nullptr, have_prev_adj ? pajy : nullptr, ajx,
                                     ajy, ajc);
                tl.first += ta.first;
                tl.second += ta.second;
                rx = ajx;
                ry = ajy;
                rc = ajc;
            }
            compose(upx, upy, rx, ry, w, h, cax, cay, s);
            conf_at(rc, upx, upy, w, h, caa, s);
            hyb_carried(cax, cay, upx, upy, ax, ay, n, s);
            // the adjacent leg's raw field starts the next adjacent leg
            dcopy(pajx, rx, n * sizeof(double), s);
            dcopy(pajy, ry, n * sizeof(double), s);
            have_adj = true;
        } else {
            dcopy(ax, upx, n * sizeof(double), s);
            dcopy(ay, upy, n * sizeof(double), s);
            hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, caa, 0.0, n);
            LAUNCHCHECK();
        }
        hyb_has_fb(flags, caa, c.conf_min, dm, cnt, n, s);
        const uint8_t *g0 = nullptr;
        if (gate_on && counters()[HYB_ANY_RE_NOFB]) {
            hyb_gate(flags, dm, cx, cy, cx, cy, w, h, radius, c.fb_min_support, c.fb_tol, gscr, gate0, s);
            g0 = gate0;
        }
        hyb_reacquire(flags, g0, cx, cy, ax, ay, legc, caa, c.fb_tol, c.conf_min, c.fb_max, c0r, anch, nfb, dm, cnt, n,
                      s);
        const uint8_t *g1 = nullptr;
        if (gate_on && counters()[HYB_ANY_CAND]) {
            hyb_gate(flags, dm, cx, cy, ax, ay, w, h, radius, c.fb_min_support, c.fb_tol, gscr, gate, s);
            g1 = gate;
        }
        Staged st = outputs(out, device_out);
        hyb_output(flags, g1, cx, cy, ax, ay, link, caa, nfb, c.fb_max, dm, hux, huy, (float *)ws.bconf.p, hlow,
                   dfb, cnt, n, s);
        HIPCHECK(hipEventRecord(ws.ev[2], s));
        finish_device(ws, hux, huy, (const float *)ws.bconf.p, hlow, w, h, dm, c, &st.d, s);
        to_f32(hux, st.d.raw_ux, n, s);
        to_f32(huy, st.d.raw_uy, n, s);
        if (st.d.conf) dcopy(st.d.conf, ws.bconf.p, n * sizeof(float), s);
        ws.prof.mark(ST_END, s);
        HIPCHECK(hipEventRecord(ws.ev[3], s));
        if (out->fb) {
            if (device_out)
                dcopy(out->fb, dfb, n * sizeof(int32_t), s);
            else
                copy_d2h(ws.pin, out->fb, dfb, n * sizeof(int32_t), s);
        }
        deliver(st, out, device_out);
        ws.prof.flush();
        const std::vector<int> hc = counters();
        const double nroi = (double)std::max(hc[HYB_N_ROI], 1);
        out->t_coarse = tl.first;
        out->t_fine = tl.second;
        out->t_strain = elapsed(ws.ev[2], ws.ev[3]) * 1e-3;
        out->reference = ref_idx;
        out->mean_conf = cm;
        out->n_fb_max = hc[HYB_NFB_MAX];
        out->fallback_frac = hc[HYB_N_FB_OK] / nroi;
        out->fb_gate_rejected = hc[HYB_N_REJECTED] / nroi;
        out->reacquired = hc[HYB_N_RE] / nroi;
        // state for frame k + 1
        dcopy(upx, hux, n * sizeof(double), s);
        dcopy(upy, huy, n * sizeof(double), s);
        hyb_good_of(flags, good_prev, n, s);
        have_prev_adj = have_adj;
        dcopy(px, lx, n * sizeof(double), s);
        dcopy(py, ly, n * sizeof(double), s);
        dcopy(pc, lc, n * sizeof(double), s);
        dcopy(fprev, fcur, n * sizeof(float), s);
        have_prev = true;
        wait_stream(s);
    }
};
}  // namespace

}  // namespace ofx

using namespace ofx;

// ---------------------------------------------------------------- allocation-free entry point
// The scratch of ofx_median_device_noalloc: one fixed array per GPU, in the library's code object,
// held until the process ends. Host code never uses its name as an address (on the host it names
// a stand-in, not device memory): each device's address comes from hipGetSymbolAddress, taken once
// per device at set-up (ofx_create) with that device current, and only that address reaches a
// kernel. Every call on a GPU shares its array, so calls that overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16384;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_noalloc_mu;
std::vector<void *> g_noalloc_ptr;  // per device; nullptr until that device's set-up

// with `device` current
void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc_ptr.size() <= device) g_noalloc_ptr.resize(device + 1, nullptr);
    if (g_noalloc_ptr[device]) return;
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(OFX_EDEVICE, "hipGetSymbolAddress gave no address for the scratch array");
    g_noalloc_ptr[device] = p;
}

void *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (device < 0 || device >= (int)g_noalloc_ptr.size() || !g_noalloc_ptr[device])
        throw Error(OFX_EDEVICE, "the scratch array of device " + std::to_string(device) + " was never set up");
    return g_noalloc_ptr[device];
}

// The selection's digit histograms count in 32 bits (SelectState in fine.hip).
void check_median_count(size_t n) {
    if (n > 0xFFFFFFFFull)
        throw Error(OFX_ETOOLARGE, "median of " + std::to_string(n) + " values: at most 4294967295 are supported");
}
}  // namespace

extern "C" {

void ofx_config_default(ofx_config *c) {
    std::memset(c, 0, sizeof(*c));
    c->grid_step = 16;
    c->poc_win = 64;
    c->max_disp = 32.0;
    c->poc_peak_min = 0.1;
    c->median_thresh = 2.0;
    c->lam = 8.0;
    c->n_scales = 3;
    c->scales[0] = 0.6;
    c->scales[1] = 1.0;
    c->scales[2] = 1.6;
    c->n_orient = 4;
    c->bandwidth = 0.5;
    c->dc_cut = 2.0;
    c->sigma_w = 4.0;
    c->shape_order = 1;
    c->sigma_s = 4.0;
    c->outer_iters = 8;
    c->tile = 0;
    c->inner_iters = 2;
    c->conf_min = 0.2;
    c->edge_guard = -1.0;
    c->edge_guard_min_area = -1.0;
    c->phase_tol = 0.4;
    c->strain_win = 15;
    c->ref_update_conf = 0.5;
    c->seq_hybrid = 0;
    c->fb_max = 10;
    c->fb_radius = -1.0;
    c->fb_min_support = 20;
    c->fb_tol = 0.5;
    c->prewarp_freeze_at = 0;
    c->coarse_learned = 0;
    c->coarse_gate = 0.9;
}

const char *ofx_last_error(void) { return g_last_error.c_str(); }

int ofx_create(int device, int lanes, ofx_engine **out) {
    return guarded([&] {
        if (!out) throw Error(OFX_EINVAL, "out is NULL");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error(OFX_ENODEVICE, "no HIP device found");
        if (device < 0 || device >= count) throw Error(OFX_ENODEVICE, "HIP device " + std::to_string(device) + " does not exist");
        if (lanes < 1 || lanes > 8) throw Error(OFX_EINVAL, "lanes must lie in [1, 8]");
        watchdog_start();
        HIPCHECK(hipSetDevice(device));
        noalloc_setup(device);
        std::unique_ptr<ofx_engine> e(new ofx_engine);
        e->device = device;
        for (int i = 0; i < lanes; i++) {
            std::unique_ptr<Lane> l(new Lane);
            HIPCHECK(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
            e->lanes.push_back(std::move(l));
        }
        *out = e.release();
    });
}

void ofx_destroy(ofx_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (auto &l : e->lanes) {
        if (l->stream) (void)hipStreamSynchronize(l->stream);
        l->ws.reset();
        if (l->stream) (void)hipStreamDestroy(l->stream);
    }
    delete e;
}

int ofx_lanes(const ofx_engine *e) { return e ? (int)e->lanes.size() : 0; }

int ofx_profile(ofx_engine *e, int enable) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        for (auto &l : e->lanes) {
            if (!l->ws) l->ws.reset(new Workspace);
            l->ws->prof.on = enable != 0;
            l->ws->prof.reset();
        }
    });
}

int ofx_profile_read(ofx_engine *e, double *ms, int64_t *calls, int n) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        for (int i = 0; i < n; i++) {
            if (ms) ms[i] = 0;
            if (calls) calls[i] = 0;
        }
        for (auto &l : e->lanes) {
            if (!l->ws) continue;
            for (int i = 0; i < n && i < ST_COUNT; i++) {
                if (ms) ms[i] += l->ws->prof.ms[i];
                if (calls) calls[i] += l->ws->prof.calls[i];
            }
        }
    });
}

const char *ofx_profile_stage(int i) { return (i >= 0 && i < ST_COUNT) ? STAGE_NAMES[i] : nullptr; }

int ofx_returns_slope(const ofx_engine *) { return 0; }

int ofx_check_window(ofx_engine *e, int w, int h, const ofx_config *cfg) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        validate(cfg);
        check_window_impl(w, h, *cfg);
        // device memory of one lane's workspace at this size, against what is free
        const size_t n = (size_t)w * h, np = (size_t)(w + 2 * PAD) * (h + 2 * PAD);
        const size_t F = (size_t)cfg->n_orient * cfg->n_scales;
        const size_t need = np * 8 * (std::max<size_t>(F, 3) + 2) + n * (16 * F + 8 * 60 + 4 * 20);
        HIPCHECK(hipSetDevice(e->device));
        size_t free_b = 0, total = 0;
        HIPCHECK(hipMemGetInfo(&free_b, &total));
        if (need > free_b)
            throw Error(OFX_ETOOLARGE, "window " + std::to_string(w) + "x" + std::to_string(h) + " needs about " +
                                           std::to_string(need >> 20) + " MiB of device memory, " +
                                           std::to_string(free_b >> 20) + " MiB free");
    });
}

int ofx_release_caches(ofx_engine *e) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        HIPCHECK(hipSetDevice(e->device));
        for (auto &l : e->lanes) {
            bool expected = false;
            if (!l->busy.compare_exchange_strong(expected, true)) continue;  // in use: keep
            wait_stream(l->stream);
            l->ws.reset();
            l->busy.store(false);
        }
    });
}

int ofx_refine_raw(ofx_engine *e, const float *i1, int w, int h, const float *i2, int w2, int h2, int64_t shift_x,
                   int64_t shift_y, const float *ux, const float *uy, const ofx_config *cfg, double *out_ux,
                   double *out_uy, double *out_energy, double *out_resid) {
    return guarded([&] {
        validate(cfg);
        if (!i1 || !i2 || !ux || !uy) throw Error(OFX_EINVAL, "input pointer is NULL");
        if (w2 < 1 || h2 < 1) throw Error(OFX_EINVAL, "refine: empty image i2");
        if (w2 > MAX_I2_SIDE || h2 > MAX_I2_SIDE)
            throw Error(OFX_ETOOLARGE, "image 2 window " + std::to_string(w2) + "x" + std::to_string(h2) +
                                           " exceeds the " + std::to_string(MAX_I2_SIDE) + " px side the spline rows hold");
        check_window_impl(w, h, *cfg);
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        const size_t n = (size_t)w * h, n2 = (size_t)w2 * h2;
        float *d1 = B<float>(ws.bi1, n), *d2 = B<float>(ws.bi2, n2);
        float *dx = B<float>(ws.binx, n), *dy = B<float>(ws.biny, n);
        copy_h2d(ws.pin, d1, i1, n * sizeof(float), s);
        copy_h2d(ws.pin, d2, i2, n2 * sizeof(float), s);
        copy_h2d(ws.pin, dx, ux, n * sizeof(float), s);
        copy_h2d(ws.pin, dy, uy, n * sizeof(float), s);
        to_f64(dx, B<double>(ws.bux, n), n, s);
        to_f64(dy, B<double>(ws.buy, n), n, s);
        refine_graph(ws, d1, d2, w, h, w2, h2, (long)shift_x, (long)shift_y, *cfg, s, false);
        if (out_ux) copy_d2h(ws.pin, out_ux, ws.bux.p, n * sizeof(double), s);
        if (out_uy) copy_d2h(ws.pin, out_uy, ws.buy.p, n * sizeof(double), s);
        if (out_energy) copy_d2h(ws.pin, out_energy, ws.benergy.p, n * sizeof(double), s);
        if (out_resid) copy_d2h(ws.pin, out_resid, ws.bresid.p, n * sizeof(double), s);
        wait_stream(s);
    });
}

int ofx_coarse_flow(ofx_engine *e, const float *i1, const float *i2, int w, int h, const float *init_ux,
                    const float *init_uy, const ofx_config *cfg, double *out_ux, double *out_uy) {
    return guarded([&] {
        validate(cfg);
        if (!i1 || !i2 || !out_ux || !out_uy) throw Error(OFX_EINVAL, "input or output pointer is NULL");
        if ((init_ux == nullptr) != (init_uy == nullptr)) throw Error(OFX_EINVAL, "init needs both components");
        check_window_impl(w, h, *cfg);
        if (std::min(w, h) < cfg->poc_win)
            throw Error(OFX_EINVAL, "image smaller than poc_win (" + std::to_string(cfg->poc_win) + ")");
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        const size_t n = (size_t)w * h;
        float *d1 = B<float>(ws.bi1, n), *d2 = B<float>(ws.bi2, n);
        copy_h2d(ws.pin, d1, i1, n * sizeof(float), s);
        copy_h2d(ws.pin, d2, i2, n * sizeof(float), s);
        double *ix = nullptr, *iy = nullptr;
        if (init_ux) {
            float *fx = B<float>(ws.binx, n), *fy = B<float>(ws.biny, n);
            copy_h2d(ws.pin, fx, init_ux, n * sizeof(float), s);
            copy_h2d(ws.pin, fy, init_uy, n * sizeof(float), s);
            ix = B<double>(ws.bstart_x, n);
            iy = B<double>(ws.bstart_y, n);
            to_f64(fx, ix, n, s);
            to_f64(fy, iy, n, s);
        }
        double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
        coarse_dense(d1, d2, w, h, *cfg, ix, iy, ux, uy, ws.coarse, s);
        copy_d2h(ws.pin, out_ux, ux, n * sizeof(double), s);
        copy_d2h(ws.pin, out_uy, uy, n * sizeof(double), s);
        wait_stream(s);
    });
}

// tavg.rs:118-163 on the device: `other` (device, w x h) resampled onto the grid of `centre` by
// the affine fit to the coarse grid (align 1, fit on the host like the reference's) or the dense
// coarse field (align 2), into `out`; false when no motion can be estimated.
bool align_device(Workspace &ws, const float *centre, const float *other, int w, int h, int align,
                  const uint8_t *mask, const ofx_config &c, float *out, hipStream_t s) {
    if (std::min(w, h) < c.poc_win) return false;  // no POC grid fits (tavg.rs:290-296)
    const size_t n = (size_t)w * h;
    double *ux = B<double>(ws.bux, n), *uy = B<double>(ws.buy, n);
    float *fx = B<float>(ws.btav_fx, n), *fy = B<float>(ws.btav_fy, n);
    coarse_dense(centre, other, w, h, c, nullptr, nullptr, ux, uy, ws.coarse, s);
    if (align == 2) {
        f64_to_f32(ux, fx, n, s);
        f64_to_f32(uy, fy, n, s);
    } else {
        std::vector<int> gxs, gys;
        std::vector<double> gux, guy, gpk, xs, ys, vx, vy;
        coarse_grid(ws.coarse, w, h, c, gxs, gys, gux, guy, gpk, ws.pin, s);
        for (size_t iy = 0; iy < gys.size(); iy++)
            for (size_t ix = 0; ix < gxs.size(); ix++) {
                const size_t j = iy * gxs.size() + ix;
                // the reference's grid is f32 (poc.rs Grid)
                if ((float)gpk[j] > 0.0f && (!mask || mask[(size_t)gys[iy] * w + gxs[ix]])) {
                    xs.push_back(gxs[ix]);
                    ys.push_back(gys[iy]);
                    vx.push_back((double)(float)gux[j]);
                    vy.push_back((double)(float)guy[j]);
                }
            }
        double cx[3], cy[3];
        if (xs.size() < 6 || !fit_affine(xs, ys, vx, vy, cx, cy)) return false;
        affine_field(cx, cy, w, h, fx, fy, s);
    }
    float *coef = B<float>(ws.btav_c, n), *tmp = B<float>(ws.btav_tmp, n);
    dcopy(coef, other, n * sizeof(float), s);
    prefilter_real(coef, tmp, w, h, s);
    warp(coef, w, h, 0, 0, fx, fy, out, w, h, s);
    return true;
}

int ofx_tavg_window(ofx_engine *e, const float *frames, int nf, int centre, int w, int h, int align,
                    const uint8_t *mask, int keep_unaligned, const ofx_config *cfg, float *out, int *n_used) {
    return guarded([&] {
        if (!frames || !out) throw Error(OFX_EINVAL, "frames or out is NULL");
        if (nf < 1 || centre < 0 || centre >= nf) throw Error(OFX_EINVAL, "centre outside the window");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty window");
        if (align < 0 || align > 2) throw Error(OFX_EINVAL, "align must be 0 (none), 1 (affine) or 2 (dense)");
        if (align) {
            if (!cfg) throw Error(OFX_EINVAL, "aligning needs a configuration");
            validate(cfg);
            if (std::min(w, h) >= cfg->poc_win) check_window_impl(w, h, *cfg);
            // the aligned frames are resampled through a B-spline of the whole frame
            if (w > MAX_I2_SIDE || h > MAX_I2_SIDE)
                throw Error(OFX_ETOOLARGE, "aligning frames of " + std::to_string(w) + "x" + std::to_string(h) +
                                               " exceeds the " + std::to_string(MAX_I2_SIDE) +
                                               " px side the spline rows hold");
        }
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        const size_t n = (size_t)w * h;
        float *dc = B<float>(ws.bi1, n), *dother = B<float>(ws.bi2, n), *da = B<float>(ws.btav_a, n);
        double *acc = B<double>(ws.btav_acc, n);
        copy_h2d(ws.pin, dc, frames + (size_t)centre * n, n * sizeof(float), s);
        mean_add(acc, dc, n, true, s);
        int used = 1;
        for (int i = 0; i < nf; i++) {
            if (i == centre) continue;
            copy_h2d(ws.pin, dother, frames + (size_t)i * n, n * sizeof(float), s);
            if (!align) {
                mean_add(acc, dother, n, false, s);
                used++;
            } else if (align_device(ws, dc, dother, w, h, align, mask, *cfg, da, s)) {
                mean_add(acc, da, n, false, s);
                used++;
            } else if (keep_unaligned) {
                mean_add(acc, dother, n, false, s);
                used++;
            }
        }
        if (used == 1)
            copy_d2h(ws.pin, out, dc, n * sizeof(float), s);
        else {
            float *dout = B<float>(ws.btav_fx, n);
            mean_finish(acc, (double)used, dout, n, s);
            copy_d2h(ws.pin, out, dout, n * sizeof(float), s);
        }
        wait_stream(s);
        if (n_used) *n_used = used;
    });
}

int ofx_frame_to_gray(ofx_engine *e, const void *pixels, int bits, int channels, int w, int h, int factor,
                      float *out) {
    return guarded([&] {
        if (!pixels || !out) throw Error(OFX_EINVAL, "pixels or out is NULL");
        if (bits != 8 && bits != 16) throw Error(OFX_EINVAL, "pixels must be 8 or 16 bit");
        if (channels != 1 && channels != 3 && channels != 4) throw Error(OFX_EINVAL, "channels must be 1, 3 or 4");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty frame");
        if (factor < 1) throw Error(OFX_EINVAL, "scale must be >= 1, got " + std::to_string(factor));
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        const size_t bytes = (size_t)w * h * channels * (bits / 8);
        const int ow = w / factor, oh = h / factor;
        if (ow < 1 || oh < 1) return;
        char *dp = B<char>(ws.bpix, bytes);
        float *dout = B<float>(ws.btav_fx, (size_t)ow * oh);
        copy_h2d(ws.pin, dp, pixels, bytes, s);
        frame_to_gray(dp, bits, channels, w, h, factor, dout, s);
        copy_d2h(ws.pin, out, dout, (size_t)ow * oh * sizeof(float), s);
        wait_stream(s);
    });
}

int ofx_set_schedule(ofx_engine *e, const ofx_schedule *sc) {
    return guarded([&] {
        if (!e || !sc) throw Error(OFX_EINVAL, "engine or schedule is NULL");
        if (sc->tile_groups < 1 || sc->tile_groups > MAX_TILE_GROUPS)
            throw Error(OFX_EINVAL, "tile_groups must lie in [1, " + std::to_string(MAX_TILE_GROUPS) + "]");
        e->tile_groups.store(sc->tile_groups);
        e->graphs.store(sc->graphs != 0);
    });
}

int ofx_get_schedule(const ofx_engine *e, ofx_schedule *sc) {
    return guarded([&] {
        if (!e || !sc) throw Error(OFX_EINVAL, "engine or schedule is NULL");
        sc->tile_groups = e->tile_groups.load();
        sc->graphs = e->graphs.load() ? 1 : 0;
    });
}

int ofx_graph_stats(ofx_engine *e, int64_t *captures, int64_t *launches) {
    return guarded([&] {
        if (!e || !captures || !launches) throw Error(OFX_EINVAL, "engine or output pointer is NULL");
        *captures = *launches = 0;
        for (auto &l : e->lanes)
            if (l->ws) {
                *captures += l->ws->graph_captures;
                *launches += l->ws->graph_launches;
            }
    });
}

size_t ofx_coarse_net_params(void) { return net_param_count(); }

int ofx_coarse_net_load(ofx_engine *e, const float *params, size_t n) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        if (!params) throw Error(OFX_EINVAL, "params is NULL");
        if (n != net_param_count())
            throw Error(OFX_EINVAL, "the coarse network has " + std::to_string(net_param_count()) + " parameters, got " +
                                        std::to_string(n));
        for (size_t i = 0; i < n; i++)
            if (!std::isfinite(params[i])) throw Error(OFX_EINVAL, "parameter " + std::to_string(i) + " is not finite");
        // every lane held while the network changes
        std::vector<Lane *> held;
        struct Release {
            std::vector<Lane *> &h;
            ~Release() {
                for (Lane *l : h) l->busy.store(false);
            }
        } release{held};
        for (auto &l : e->lanes) {
            bool expected = false;
            if (!l->busy.compare_exchange_strong(expected, true))
                throw Error(OFX_EBUSY, "a call is running on the engine: the network cannot change now");
            held.push_back(l.get());
        }
        HIPCHECK(hipSetDevice(e->device));
        std::vector<float> m;
        net_flat_to_master(params, m);
        if (!e->net) e->net.reset(new CoarseNet);
        e->net->upload(m, e->lanes[0]->stream);
        for (auto &l : e->lanes)
            if (l->ws) l->ws->net = e->net.get();
    });
}

int ofx_coarse_net_propose(ofx_engine *e, const float *i1, const float *i2, int w, int h, float *out) {
    return guarded([&] {
        if (!i1 || !i2 || !out) throw Error(OFX_EINVAL, "input or output pointer is NULL");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty frame");
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        if (!ws.net) throw Error(OFX_EINVAL, "no coarse network loaded (ofx_coarse_net_load)");
        const size_t n = (size_t)w * h;
        float *d1 = B<float>(ws.bi1, n), *d2 = B<float>(ws.bi2, n);
        copy_h2d(ws.pin, d1, i1, n * sizeof(float), s);
        copy_h2d(ws.pin, d2, i2, n * sizeof(float), s);
        if (!ws.nacts) ws.nacts.reset(new NetActs);
        net_propose(*ws.net, *ws.nacts, d1, d2, w, h, s);
        const size_t nc = (size_t)ws.nacts->w8 * ws.nacts->h8 * 2;
        copy_d2h(ws.pin, out, ws.nacts->pred, nc * sizeof(float), s);
        wait_stream(s);
        for (size_t i = 0; i < nc; i++) out[i] *= 8.f;
    });
}

int ofx_coarse_net_gradient(ofx_engine *e, const float *params, size_t n, const float *a1, const float *a2, int nb,
                            int w8, int h8, const float *target, double *loss, float *grad, float *pred) {
    return guarded([&] {
        if (!params || !a1 || !a2 || !target || !loss || !grad) throw Error(OFX_EINVAL, "input or output pointer is NULL");
        if (n != net_param_count()) throw Error(OFX_EINVAL, "wrong number of parameters");
        if (nb < 1 || w8 < 1 || h8 < 1) throw Error(OFX_EINVAL, "empty batch");
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        std::vector<float> m;
        net_flat_to_master(params, m);
        CoarseNet net;
        net.upload(m, s);
        NetTrainer tr;
        tr.ensure(nb, h8, w8);
        const size_t nc = (size_t)nb * w8 * h8;
        copy_h2d(ws.pin, tr.A.a1, a1, nc * sizeof(float), s);
        copy_h2d(ws.pin, tr.A.a2, a2, nc * sizeof(float), s);
        copy_h2d(ws.pin, tr.target, target, 2 * nc * sizeof(float), s);
        GrowBuf lb;
        double *dl = lb.get<double>(1);
        tr.forward_backward(net, dl, s);
        copy_d2h(ws.pin, loss, dl, sizeof(double), s);
        std::vector<float> gm(net_master_floats());
        copy_d2h(ws.pin, gm.data(), tr.grad, gm.size() * sizeof(float), s);
        if (pred) copy_d2h(ws.pin, pred, tr.A.pred, 2 * nc * sizeof(float), s);
        wait_stream(s);
        net_master_to_flat(gm, grad);
    });
}

int ofx_coarse_net_train(ofx_engine *e, const ofx_net_train *t, const float *patterns, int n_pat, int pw, int ph,
                         float *params, double *loss) {
    return guarded([&] {
        if (!t || !patterns || !params) throw Error(OFX_EINVAL, "input or output pointer is NULL");
        if (t->steps < 1 || t->batch < 1 || t->n_pairs < t->batch || t->n_pairs % t->batch)
            throw Error(OFX_EINVAL, "steps >= 1, batch >= 1 and n_pairs a positive multiple of batch required");
        if (t->crop < 64 || t->crop % 8) throw Error(OFX_EINVAL, "crop must be a multiple of 8 and >= 64");
        if (!(t->lr > 0) || !(t->lr_final >= 0) || !(t->max_disp >= 0) || !(t->noise >= 0))
            throw Error(OFX_EINVAL, "lr > 0, lr_final >= 0, max_disp >= 0 and noise >= 0 required");
        if (n_pat < 1 || pw < t->crop + 129 || ph < t->crop + 129)
            throw Error(OFX_EINVAL, "patterns must exceed the crop by more than 128 px a side");
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        step_mark("coarse net: setting up the training");
        std::vector<float> m;
        net_flat_to_master(params, m);
        CoarseNet net;
        net.upload(m, s);
        GrowBuf pb, sb, lb;
        float *pats = pb.get<float>((size_t)n_pat * pw * ph);
        copy_h2d(ws.pin, pats, patterns, (size_t)n_pat * pw * ph * sizeof(float), s);
        std::vector<TrainSample> smp;
        make_train_samples(t->seed, 0, t->n_pairs, n_pat, pw, ph, t->crop, t->max_disp, smp);
        TrainSample *dsmp = sb.get<TrainSample>(smp.size());
        copy_h2d(ws.pin, dsmp, smp.data(), smp.size() * sizeof(TrainSample), s);
        double *dloss = lb.get<double>(t->steps);
        const int c8 = t->crop / 8;
        NetTrainer tr;
        tr.ensure(t->batch, c8, c8);
        tr.zero_moments(s);
        step_mark("coarse net: training steps");
        for (int k = 0; k < t->steps; k++) {
            const int first = (int)(((long)k * t->batch) % t->n_pairs);
            render_train_pairs(pats, pw, ph, dsmp + first, t->batch, c8, c8, (float)t->noise, tr.A.a1, tr.A.a2,
                               tr.target, s);
            tr.forward_backward(net, dloss + k, s);
            const double f = t->steps > 1 ? (double)k / (t->steps - 1) : 1.0;
            const double lr = t->lr_final + 0.5 * (t->lr - t->lr_final) * (1.0 + std::cos(M_PI * f));
            tr.adam_step(net, k + 1, lr, s);
        }
        step_mark("coarse net: reading the trained parameters");
        std::vector<float> out(net_master_floats());
        copy_d2h(ws.pin, out.data(), net.master, out.size() * sizeof(float), s);
        if (loss) copy_d2h(ws.pin, loss, dloss, (size_t)t->steps * sizeof(double), s);
        wait_stream(s);
        net_master_to_flat(out, params);
    });
}

int ofx_median(ofx_engine *e, const double *values, size_t n, double *out) {
    return guarded([&] {
        if (!out || (n && !values)) throw Error(OFX_EINVAL, "values or out is NULL");
        check_median_count(n);
        if (n == 0) {
            *out = NAN;  // numpy.median of nothing
            return;
        }
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        double *dv = B<double>(ws.benergy, n);
        double *med = B<double>(ws.bmed, 4);
        copy_h2d(ws.pin, dv, values, n * sizeof(double), s);
        median(dv, n, ws.bmedscr.get<char>(median_scratch_bytes()), med, s);
        copy_d2h(ws.pin, out, med, sizeof(double), s);
        wait_stream(s);
    });
}

int ofx_median_device_noalloc(ofx_engine *e, const double *d_values, size_t n, double *d_out) {
    return guarded([&] {
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        if (!d_out || (n && !d_values)) throw Error(OFX_EINVAL, "d_values or d_out is NULL");
        check_median_count(n);
        static_assert(NOALLOC_SCRATCH_BYTES >= sizeof(double), "scratch too small");
        void *scr = noalloc_scratch(e->device);
        if (median_scratch_bytes() > NOALLOC_SCRATCH_BYTES)
            throw Error(OFX_EUNSUPPORTED, "the selection state exceeds the library's fixed scratch array");
        LaneGuard g(e);
        hipStream_t s = g.lane->stream;
        if (n == 0) {
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, d_out, (double)NAN, (size_t)1);  // numpy.median of nothing
            LAUNCHCHECK();
        } else {
            median(d_values, n, scr, d_out, s);
        }
        wait_stream(s);
    });
}

int ofx_track_pair_device(ofx_engine *e, const float *d_i1, const float *d_i2, int w, int h, const uint8_t *d_mask,
                          const float *d_init_ux, const float *d_init_uy, const ofx_config *cfg, ofx_result *d_out) {
    return guarded([&] {
        validate(cfg);
        if (!d_i1 || !d_i2 || !d_out) throw Error(OFX_EINVAL, "input or output pointer is NULL");
        if ((d_init_ux == nullptr) != (d_init_uy == nullptr)) throw Error(OFX_EINVAL, "init needs both components");
        check_window_impl(w, h, *cfg);
        if (std::min(w, h) < cfg->poc_win)
            throw Error(OFX_EINVAL, "image smaller than poc_win (" + std::to_string(cfg->poc_win) + ")");
        LaneGuard g(e);
        Workspace &ws = *g.lane->ws;
        hipStream_t s = g.lane->stream;
        const size_t n = (size_t)w * h;
        double *ix = nullptr, *iy = nullptr;
        if (d_init_ux) {
            ix = B<double>(ws.bstart_x, n);
            iy = B<double>(ws.bstart_y, n);
            to_f64(d_init_ux, ix, n, s);
            to_f64(d_init_uy, iy, n, s);
        }
        pair_device(ws, d_i1, d_i2, w, h, d_mask, ix, iy, *cfg, d_out, s);
        record_times(ws, d_out);
    });
}

int ofx_track_pair(ofx_engine *e, const float *i1, const float *i2, int w, int h, const uint8_t *mask,
                   const float *init_ux, const float *init_uy, const ofx_config *cfg, ofx_result *out) {
    return guarded([&] {
        validate(cfg);
        check_pair_args(i1, i2, w, h, mask, init_ux, init_uy, *cfg, out);
        LaneGuard g(e);
        track_pair_on_lane(*g.lane, i1, i2, w, h, mask, init_ux, init_uy, *cfg, out);
    });
}

// Independent pairs over the engine's free lanes (the batched workload of many uneven pairs):
// one host thread per lane pulls the next pair, largest first, and runs it on the lane's stream
// exactly as ofx_track_pair would, so each result is bit-identical to the single call.
int ofx_track_pairs(ofx_engine *e, int n, const ofx_pair *pairs, const ofx_config *cfg, ofx_result *results) {
    return guarded([&] {
        validate(cfg);
        if (n < 0 || (n > 0 && (!pairs || !results))) throw Error(OFX_EINVAL, "pairs or results is NULL");
        for (int i = 0; i < n; i++) {
            const ofx_pair &p = pairs[i];
            try {
                check_pair_args(p.i1, p.i2, p.w, p.h, p.mask, p.init_ux, p.init_uy, *cfg, &results[i]);
            } catch (const Error &er) {
                throw Error(er.code, "pair " + std::to_string(i) + ": " + er.what());
            }
        }
        if (!e) throw Error(OFX_EINVAL, "engine is NULL");
        if (n == 0) return;
        std::vector<Lane *> held;
        for (auto &l : e->lanes) {
            bool expected = false;
            if (l->busy.compare_exchange_strong(expected, true)) held.push_back(l.get());
        }
        if (held.empty()) throw Error(OFX_EBUSY, "every lane of the GPU engine is busy");
        struct Release {
            std::vector<Lane *> &h;
            ~Release() {
                for (Lane *l : h) l->busy.store(false);
            }
        } release{held};
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return (long long)pairs[a].w * pairs[a].h > (long long)pairs[b].w * pairs[b].h;
        });
        std::atomic<int> next{0};
        std::vector<int> codes(n, OFX_OK);
        std::vector<std::string> msgs(n);
        // memory a growing buffer gives up is freed after every lane has stopped
        std::mutex held_mu;
        std::vector<void *> given_up;
        auto work = [&](Lane *lane) {
            Progress &pg = progress();
            pg.call_id.fetch_add(1);
            pg.call_since_ms.store(now_ms());
            struct Clear {
                Progress &p;
                ~Clear() { p.call_since_ms.store(0); }
            } clear{pg};
            if (hipSetDevice(e->device) != hipSuccess) {
                for (int k; (k = next.fetch_add(1)) < n;) {
                    codes[order[k]] = OFX_EDEVICE;
                    msgs[order[k]] = "hipSetDevice failed";
                }
                return;
            }
            if (!lane->ws) lane->ws.reset(new Workspace);
            DeferFrees defer;
            struct Hand {
                DeferFrees &d;
                std::mutex &mu;
                std::vector<void *> &out;
                ~Hand() {
                    std::lock_guard<std::mutex> lk(mu);
                    out.insert(out.end(), d.held.begin(), d.held.end());
                    d.held.clear();
                }
            } hand{defer, held_mu, given_up};
            for (int k; (k = next.fetch_add(1)) < n;) {
                const int i = order[k];
                const ofx_pair &p = pairs[i];
                try {
                    track_pair_on_lane(*lane, p.i1, p.i2, p.w, p.h, p.mask, p.init_ux, p.init_uy, *cfg, &results[i]);
                } catch (const Error &er) {
                    codes[i] = er.code;
                    msgs[i] = er.what();
                } catch (const std::exception &ex) {
                    codes[i] = OFX_EDEVICE;
                    msgs[i] = ex.what();
                }
            }
        };
        std::vector<std::thread> threads;
        try {
            for (size_t t = 1; t < held.size(); t++) threads.emplace_back(work, held[t]);
        } catch (const std::exception &) {
            // no thread for the remaining lanes: the running workers take their pairs
        }
        work(held[0]);
        for (auto &t : threads) t.join();
        for (void *p : given_up) (void)hipFree(p);
        for (int i = 0; i < n; i++)
            if (codes[i] != OFX_OK) throw Error(codes[i], "pair " + std::to_string(i) + ": " + msgs[i]);
    });
}

// track.py:260-315 / track/hybrid.rs:262-540 over a whole frame stack: begin, one step per
// frame, end, on one lane.
int ofx_track_sequence(ofx_engine *e, const float *frames, int nf, int w, int h, const uint8_t *mask,
                       const ofx_config *cfg, ofx_result *results) {
    return guarded([&] {
        validate(cfg);
        if (!frames || !results) throw Error(OFX_EINVAL, "frames or results is NULL");
        if (nf < 1) throw Error(OFX_EINVAL, "no frames");
        check_sequence_args(w, h, mask, *cfg);
        LaneGuard g(e);
        Seq q(e, g.lane, *cfg, w, h);
        q.begin(frames, false, mask);
        for (int k = 1; k < nf; k++) q.step(frames + (size_t)k * q.n, false, &results[k - 1], false);
    });
}

// The streaming form: the sequence holds its lane from ofx_seq_begin to ofx_seq_end.
struct ofx_sequence {
    ofx_engine *e = nullptr;
    Lane *lane = nullptr;
    Seq *q = nullptr;
};

int ofx_seq_begin(ofx_engine *e, const float *frame0, int w, int h, const uint8_t *mask, const ofx_config *cfg,
                  ofx_sequence **out) {
    return guarded([&] {
        validate(cfg);
        if (!frame0 || !out) throw Error(OFX_EINVAL, "frame0 or out is NULL");
        *out = nullptr;
        check_sequence_args(w, h, mask, *cfg);
        LaneGuard g(e);
        std::unique_ptr<Seq> q(new Seq(e, g.lane, *cfg, w, h));
        q->begin(frame0, false, mask);
        ofx_sequence *sq = new ofx_sequence;
        sq->e = e;
        sq->lane = g.lane;
        sq->q = q.release();
        g.lane = nullptr;  // the sequence keeps the lane
        *out = sq;
    });
}

int ofx_seq_step(ofx_sequence *sq, const float *frame, ofx_result *out) {
    return guarded([&] {
        if (!sq || !sq->q) throw Error(OFX_EINVAL, "sequence is NULL or ended");
        if (!frame || !out) throw Error(OFX_EINVAL, "frame or out is NULL");
        HIPCHECK(hipSetDevice(sq->e->device));
        sq->q->step(frame, false, out, false);
    });
}

int ofx_seq_step_device(ofx_sequence *sq, const float *d_frame, ofx_result *d_out) {
    return guarded([&] {
        if (!sq || !sq->q) throw Error(OFX_EINVAL, "sequence is NULL or ended");
        if (!d_frame || !d_out) throw Error(OFX_EINVAL, "frame or out is NULL");
        HIPCHECK(hipSetDevice(sq->e->device));
        sq->q->step(d_frame, true, d_out, true);
    });
}

void ofx_seq_end(ofx_sequence *sq) {
    if (!sq) return;
    if (sq->q) {
        (void)hipSetDevice(sq->e->device);
        (void)hipStreamSynchronize(sq->lane->stream);
        delete sq->q;
    }
    if (sq->lane) sq->lane->busy.store(false);
    delete sq;
}

}  // extern "C"
