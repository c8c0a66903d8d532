// Batched 2-D FFTs of the Gabor filter stage, written for CDNA4: each workgroup holds whole
// lines in LDS (double-buffered self-sorting Stockham, mixed radix 2/3/4/5/8 plus a direct pass
// for any other prime, so every padded length W + 64 works without Bluestein). The reference
// calls rustfft on the CPU (rust/crates/optflow-core/src/fft.rs:27-116) and Metal Stockham
// shaders on the GPU (rust/crates/optflow-gpu/shaders/fft.metal:462-504).
//
// Fusions: the forward row pass builds the reflect-padded input (np.pad mode="reflect",
// phaseflow.py:92-93) on load; the inverse's first (column) pass multiplies by the Gabor filter
// G_k evaluated on the fly (phaseflow.py:73-89), optionally by i fx or i fy (the phase
// derivatives, phaseflow.py:106-107), and applies the B-spline prefilter along y; the second
// (row) pass applies it along x together with the 1/N scaling. Inverse transforms use
// conj(DFT(conj(x))) so that every pass runs the forward butterflies.
#include <cmath>
#include <map>
#include <mutex>

#include "internal.hpp"
#include "lines.hpp"

namespace ofx {

namespace {

struct PlanArgs {
    int n;
    int nfac;
    int fac[MAX_FACTORS];
};

PlanArgs args_of(const FftPlan &p) {
    PlanArgs a;
    a.n = p.n;
    a.nfac = p.nfac;
    for (int i = 0; i < MAX_FACTORS; i++) a.fac[i] = p.fac[i];
    return a;
}

__device__ __forceinline__ int reflect_idx(int i, int n) {
    if (i < 0) return -i;
    if (i >= n) return 2 * (n - 1) - i;
    return i;
}

// numpy.fft.fftfreq(n)[i] * 2 pi
__device__ __forceinline__ float freq(int i, int n) {
    int k = (i <= (n - 1) / 2) ? i : i - n;
    return (float)(6.283185307179586 * (double)k / (double)n);
}

// Forward row pass over the padded image: row y of the (h + 2 PAD) x (w + 2 PAD) reflect pad.
__global__ void __launch_bounds__(256) k_rows_fwd_reflect(const float *__restrict__ img, int w, int h,
                                                          float2 *__restrict__ out, PlanArgs P,
                                                          const float2 *__restrict__ tw) {
    extern __shared__ float2 lds[];
    const int n = P.n;
    float2 *a = lds, *b = lds + n;
    const int y = blockIdx.x;
    const float *row = img + (size_t)reflect_idx(y - PAD, h) * w;
    for (int x = threadIdx.x; x < n; x += blockDim.x) a[x] = make_float2(row[reflect_idx(x - PAD, w)], 0.f);
    __syncthreads();
    float2 *r = fft_lines_lds(a, b, n, 1, P.fac, P.nfac, tw, threadIdx.x, blockDim.x);
    float2 *o = out + (size_t)y * n;
    for (int x = threadIdx.x; x < n; x += blockDim.x) o[x] = r[x];
}

// Maps a workgroup id so that consecutive column groups run on the same XCD (workgroups are
// dispatched round-robin over the 8 XCDs); neighbours then share their L2 lines.
__device__ __forceinline__ int xcd_swizzle(int bid, int nb) {
    const int nx = 8;
    const int per = (nb + nx - 1) / nx;
    const int xcd = bid % nx, slot = bid / nx;
    const int full = nb % nx == 0 ? nx : nb % nx;  // XCDs that get `per` groups
    int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return start + slot;
}

// Column pass: CB adjacent columns per workgroup. Mode 0: plain forward transform. Mode 1: first
// pass of an inverse: loads conj(spec * G_k * mul), optionally prefilters along y.
template <int CB, int MODE>
__global__ void __launch_bounds__(256) k_cols(const float2 *__restrict__ in, float2 *__restrict__ out, int wp,
                                              PlanArgs P, const float2 *__restrict__ tw, Bank bank, int k,
                                              int mul, int prefilter) {
    extern __shared__ float2 lds[];
    const int n = P.n;  // column length hp
    float2 *a = lds, *b = lds + CB * n;
    const int ngroups = (wp + CB - 1) / CB;
    const int g = xcd_swizzle(blockIdx.x, gridDim.x);
    if (g >= ngroups) return;
    const int x0 = g * CB;
    for (int item = threadIdx.x; item < CB * n; item += blockDim.x) {
        const int y = item / CB, c = item - y * CB;
        const int x = x0 + c;
        float2 v = make_float2(0.f, 0.f);
        if (x < wp) {
            v = in[(size_t)y * wp + x];
            if (MODE == 1) {
                const float fx = freq(x, wp), fy = freq(y, n);
                const float r2 = fx * fx + fy * fy;
                const float hp = -expm1f(-r2 * bank.s_hp2_half);
                const float ax = fx - bank.w0x[k], ay = fy - bank.w0y[k];
                const float G = hp * expf(-(ax * ax + ay * ay) * bank.inv2s2[k]);
                v = cscale(v, G);
                if (mul == MUL_G_IFX) v = make_float2(-v.y * fx, v.x * fx);
                else if (mul == MUL_G_IFY) v = make_float2(-v.y * fy, v.x * fy);
                v = cconj(v);
            }
        }
        a[c * n + y] = v;
    }
    __syncthreads();
    float2 *r = fft_lines_lds(a, b, n, CB, P.fac, P.nfac, tw, threadIdx.x, blockDim.x);
    if (MODE == 1 && prefilter) {
        float2 *s = (r == a) ? b : a;
        prefilter_lines<float2>(r, s, r, n, CB, 1.0, threadIdx.x, blockDim.x);
    }
    for (int item = threadIdx.x; item < CB * n; item += blockDim.x) {
        const int y = item / CB, c = item - y * CB;
        const int x = x0 + c;
        if (x < wp) out[(size_t)y * wp + x] = r[c * n + y];
    }
}

// Second pass of an inverse: row transform, optional B-spline prefilter along x, scaling, conj.
__global__ void __launch_bounds__(256) k_rows_inv(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                  PlanArgs P, const float2 *__restrict__ tw, double scale,
                                                  int prefilter) {
    extern __shared__ float2 lds[];
    const int n = P.n;
    float2 *a = lds, *b = lds + n;
    const int y = blockIdx.x;
    const float2 *row = in + (size_t)y * n;
    for (int x = threadIdx.x; x < n; x += blockDim.x) a[x] = row[x];
    __syncthreads();
    float2 *r = fft_lines_lds(a, b, n, 1, P.fac, P.nfac, tw, threadIdx.x, blockDim.x);
    float2 *o = out + (size_t)y * n;
    if (prefilter) {
        float2 *s = (r == a) ? b : a;
        prefilter_lines<float2>(r, s, r, n, 1, scale, threadIdx.x, blockDim.x);
        for (int x = threadIdx.x; x < n; x += blockDim.x) o[x] = cconj(r[x]);
    } else {
        const float sc = (float)scale;
        for (int x = threadIdx.x; x < n; x += blockDim.x) o[x] = cconj(cscale(r[x], sc));
    }
}

// Plain forward column pass reuses k_cols<CB, 0>.

std::mutex g_plan_mu;
std::map<int, FftPlan> g_plans;

}  // namespace

const FftPlan &fft_plan(int n) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    if (n < 1 || n > 4096)
        throw Error(OFX_ETOOLARGE, "FFT length " + std::to_string(n) + " exceeds the 4096 samples an LDS line holds");
    FftPlan p;
    p.n = n;
    int m = n;
    auto push = [&](int r) {
        if (p.nfac >= MAX_FACTORS) throw Error(OFX_EUNSUPPORTED, "FFT length has too many factors");
        p.fac[p.nfac++] = r;
        m /= r;
    };
    while (m % 8 == 0) push(8);
    while (m % 4 == 0) push(4);
    while (m % 2 == 0) push(2);
    while (m % 3 == 0) push(3);
    while (m % 5 == 0) push(5);
    for (int f = 7; m > 1;) {
        if (m % f == 0) push(f);
        else f += 2;
    }
    std::vector<float2> tw(n);
    for (int e = 0; e < n; e++) {
        double a = -2.0 * M_PI * (double)e / (double)n;
        tw[e] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    HIPCHECK(hipMalloc(&p.tw, n * sizeof(float2)));
    HIPCHECK(hipMemcpy(p.tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    return g_plans[n] = p;
}

Bank make_bank(const ofx_config &c) {
    Bank b;
    b.F = c.n_orient * c.n_scales;
    double maxs = c.scales[0];
    for (int i = 1; i < c.n_scales; i++) maxs = std::max(maxs, c.scales[i]);
    const double s_hp = c.dc_cut * c.lam * maxs / (2 * M_PI);
    b.s_hp2_half = (float)(s_hp * s_hp / 2);
    int f = 0;
    for (int si = 0; si < c.n_scales; si++) {
        const double w0 = 2 * M_PI / (c.lam * c.scales[si]);
        const double s = c.bandwidth * w0;
        for (int k = 0; k < c.n_orient; k++, f++) {
            const double th = M_PI * k / c.n_orient;
            b.w0x[f] = (float)(w0 * std::cos(th));
            b.w0y[f] = (float)(w0 * std::sin(th));
            b.inv2s2[f] = (float)(1.0 / (2 * s * s));
        }
    }
    return b;
}

namespace {
int col_block(int hp) {
    // columns per workgroup: as many as keep the two LDS buffers within 64 KiB
    if (2 * 4 * hp * (int)sizeof(float2) <= 65536) return 4;
    if (2 * 2 * hp * (int)sizeof(float2) <= 65536) return 2;
    return 1;
}

template <int MODE>
void launch_cols(const float2 *in, float2 *out, int hp, int wp, const Bank *bank, int k, int mul, int prefilter,
                 hipStream_t s) {
    const FftPlan &P = fft_plan(hp);
    const int cb = col_block(hp);
    Bank dummy{};
    const Bank &bk = bank ? *bank : dummy;
    const int groups = (wp + cb - 1) / cb;
    const size_t lds = 2 * (size_t)cb * hp * sizeof(float2);
    if (cb == 4)
        hipLaunchKernelGGL((k_cols<4, MODE>), dim3(groups), dim3(256), lds, s, in, out, wp, args_of(P), P.tw, bk, k, mul,
                           prefilter);
    else if (cb == 2)
        hipLaunchKernelGGL((k_cols<2, MODE>), dim3(groups), dim3(256), lds, s, in, out, wp, args_of(P), P.tw, bk, k, mul,
                           prefilter);
    else
        hipLaunchKernelGGL((k_cols<1, MODE>), dim3(groups), dim3(256), lds, s, in, out, wp, args_of(P), P.tw, bk, k, mul,
                           prefilter);
    LAUNCHCHECK();
}
}  // namespace

void fft_forward_padded(const float *img, int w, int h, float2 *spec, float2 *tmp, hipStream_t s) {
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    const FftPlan &PR = fft_plan(wp);
    hipLaunchKernelGGL(k_rows_fwd_reflect, dim3(hp), dim3(256), 2 * wp * sizeof(float2), s, img, w, h, tmp,
                       args_of(PR), PR.tw);
    LAUNCHCHECK();
    launch_cols<0>(tmp, spec, hp, wp, nullptr, 0, 0, 0, s);
}

void fft_inverse_filtered(const float2 *spec, int hp, int wp, const Bank &bank, int k, SpecMul mul, bool prefilter,
                          float2 *tmp, float2 *out, hipStream_t s) {
    launch_cols<1>(spec, tmp, hp, wp, &bank, k, (int)mul, prefilter ? 1 : 0, s);
    const FftPlan &PR = fft_plan(wp);
    const double scale = 1.0 / ((double)hp * (double)wp);
    hipLaunchKernelGGL(k_rows_inv, dim3(hp), dim3(256), 2 * wp * sizeof(float2), s, tmp, out, args_of(PR), PR.tw,
                       scale, prefilter ? 1 : 0);
    LAUNCHCHECK();
}

void fft_release() {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    for (auto &kv : g_plans) (void)hipFree(kv.second.tw);
    g_plans.clear();
}

}  // namespace ofx
