// Internal declarations shared by the HIP translation units of libofx.so.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ofx.h"

namespace ofx {

constexpr int PAD = 32;               // reflect pad of the filter bank (phaseflow.py:47)
constexpr double APERTURE_TOL = 1e-3; // phaseflow.py:50
constexpr double RIDGE = 1e-12;       // phaseflow.py:153
constexpr int MAX_FILTERS = 64;
constexpr int MAX_FACTORS = 24;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Progress of each host thread in the library, for the watchdog (engine.hip): the stage its call
// is in, the HIP runtime call it is inside (and since when), the source line of its last kernel
// launch, and whether its last wait found work still pending on the device. A call that runs past
// OFX_WATCHDOG_S seconds (default 60) gets one report on stderr per thread, so a run that stops
// making progress says where every thread stands.
struct Progress {
    std::atomic<bool> used{false};
    std::atomic<const char *> step{"-"};
    std::atomic<const char *> hip_call{nullptr};
    std::atomic<long long> hip_since_ms{0};
    std::atomic<const char *> launch_file{"-"};
    std::atomic<int> launch_line{0};
    std::atomic<long long> call_since_ms{0};
    std::atomic<long long> call_id{0};
    std::atomic<int> pending{-1};  // last wait: 1 device work still pending, 0 done, -1 none yet
};
Progress &progress();
long long now_ms();
void step_mark(const char *what);
void launch_mark(const char *file, int line);
struct HipCallMark {
    explicit HipCallMark(const char *call);
    ~HipCallMark();
};
// hipStreamSynchronize / hipEventSynchronize by polling the stream or event, which records whether
// device work is still pending while the thread waits
void wait_stream(hipStream_t s);
void wait_event(hipEvent_t e);

// Device memory a growing buffer gives up: freed at once, or, while a deferral scope is active on
// this thread (the batched call, whose other lanes keep working), kept until the scope ends.
void release_device(void *p);
struct DeferFrees {
    DeferFrees();
    ~DeferFrees();  // frees what was deferred on this thread
    std::vector<void *> held;
    DeferFrees *prev;
};

// Copies by kernel (copy.hip): device or host-mapped pinned memory on either side, enqueued on s.
void dcopy(void *dst, const void *src, size_t bytes, hipStream_t s);
void dcopy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipStream_t s);
void make_twiddles(float2 *tw, int n, hipStream_t s);

// Host copies go through host-mapped pinned staging of at most PIN_CHUNK bytes per step, moved
// by dcopy: each call returns with the data in place.
constexpr size_t PIN_CHUNK = 8u << 20;
struct PinnedBuf {
    char *p = nullptr;   // host address
    char *dp = nullptr;  // the same memory as the device sees it
    ~PinnedBuf();
    char *get();
};
void copy_h2d(PinnedBuf &pin, void *dev, const void *host, size_t bytes, hipStream_t s);
void copy_d2h(PinnedBuf &pin, void *host, const void *dev, size_t bytes, hipStream_t s);

#define HIPCHECK(x)                                                                          \
    do {                                                                                     \
        ::ofx::HipCallMark mark_(#x);                                                        \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            throw ::ofx::Error(OFX_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHECK()                                   \
    do {                                                \
        ::ofx::launch_mark(__FILE__, __LINE__);         \
        HIPCHECK(hipGetLastError());                    \
    } while (0)

// Runs f, turning an exception into the status code it carries and the thread's last-error text
// (every C entry point goes through it).
int guarded_call(const std::function<void()> &f);
// The reference's configuration checks (config.py:166-245) for the keys the engine reads.
void validate_config(const ofx_config *cfg);
// The guard band in px: edge_guard, or with -1 the automatic 3 lam max(scales) / (2 pi bandwidth).
double guard_band_px(const ofx_config &c);
// The coarse stage's pyramid (poc.rs:610-666): the number of levels below the full image for a
// w x h image, and the POC window at a level of h x w (false: none fits).
int pyramid_levels(double max_disp, int win, int h, int w);
bool poc_window_at(int h, int w, int win, int &wy, int &wx);

// ---------------------------------------------------------------- FFT plans and passes
struct FftPlan {
    int n = 0;
    int nfac = 0;
    int fac[MAX_FACTORS] = {0};
    float2 *tw = nullptr;  // device: exp(-2 pi i e / n), e in [0, n)
    // Lengths with a prime factor of at least BLUE_MIN_PRIME are transformed as a chirp-z
    // convolution of length blue_m (0: direct passes).
    int blue_m = 0;
    float2 *chirp = nullptr;  // device: exp(i pi j^2 / n), j in [0, n)
    float2 *chirp_spec = nullptr;  // device: length-blue_m DFT of the wrapped chirp, divided by blue_m
    // Lengths with a compile-time form (16 R1 R2, fft.hip): the twiddle tables of its passes 1 and 2
    float2 *ct_tw = nullptr;
};
constexpr int BLUE_MIN_PRIME = 67;

const FftPlan &fft_plan(int n, hipStream_t s);  // cached per length, device tables never freed
void fft_release();

// Filter bank parameters for the on-the-fly evaluation of G_k (phaseflow.py:73-89).
struct Bank {
    int F = 0;
    float w0x[MAX_FILTERS], w0y[MAX_FILTERS], inv2s2[MAX_FILTERS];
    float s_hp2_half;  // s_hp^2 / 2
};
Bank make_bank(const ofx_config &c);

// What the first (column) pass of an inverse multiplies the spectrum by.
enum SpecMul { MUL_G = 0, MUL_G_IFX = 1, MUL_G_IFY = 2 };

// Forward 2-D FFT of the reflect-padded real image img (w x h, pitch w) into spec, stored
// transposed: spec[kx * hp + ky]. tmp, tmp2: hp x wp scratch.
void fft_forward_padded(const float *img, int w, int h, float2 *spec, float2 *tmp, float2 *tmp2, hipStream_t s);
// Inverse 2-D FFTs of spec * G_k (* i fx or i fy) for the K filters k = k0 .. k0 + K - 1 into the
// K planes of out (hp x wp each, natural layout), batched in one launch per pass so that the
// lines of every filter share the GPU. With prefilter, the cubic B-spline prefilter is applied
// along both axes (out holds spline coefficients). tmp, tmp2: K planes of scratch each.
void fft_inverse_filtered(const float2 *spec, int hp, int wp, const Bank &bank, int k0, int K, SpecMul mul,
                          bool prefilter, float2 *tmp, float2 *tmp2, float2 *out, hipStream_t s);

// ---------------------------------------------------------------- fine stage and friends
struct Workspace;  // per-size device buffers (engine.hip)

}  // namespace ofx
