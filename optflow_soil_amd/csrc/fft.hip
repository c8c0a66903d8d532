// Batched 2-D FFTs of the Gabor filter stage, written for CDNA4: each workgroup holds one whole
// line in LDS (double-buffered self-sorting Stockham, mixed radix 2/3/4/5/8 plus a direct pass for
// any other prime, so every padded length W + 64 works without Bluestein). The reference calls
// rustfft on the CPU (rust/crates/optflow-core/src/fft.rs:27-116) and Metal Stockham shaders on
// the GPU (rust/crates/optflow-gpu/shaders/fft.metal:462-504).
//
// Every pass transforms contiguous lines (one workgroup per line, twiddles staged in LDS); spectra
// are kept transposed ([kx][ky]) so that a 2-D transform is line pass, LDS-tiled transpose, line
// pass. Fusions: the forward pass builds the reflect-padded input (np.pad mode="reflect",
// phaseflow.py:92-93) on load; the inverse's first pass (along ky) multiplies by the Gabor filter
// G_k evaluated on the fly (phaseflow.py:73-89), optionally by i fx or i fy (the phase
// derivatives, phaseflow.py:106-107), and applies the B-spline prefilter along y (it commutes
// with the transform along x); the second pass applies it along x with the 1/N scaling. Inverse
// transforms use conj(DFT(conj(x))) so that every pass runs the forward butterflies.
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
    unsigned int magic[MAX_FACTORS];  // mod_magic of the product of the radices before pass s
};

PlanArgs args_of(const FftPlan &p) {
    PlanArgs a;
    a.n = p.n;
    a.nfac = p.nfac;
    unsigned int prod = 1;
    for (int i = 0; i < MAX_FACTORS; i++) {
        a.fac[i] = p.fac[i];
        a.magic[i] = mod_magic(prod);
        if (i < p.nfac) prod *= (unsigned int)p.fac[i];
    }
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

// The same with the step 2 pi / n given in float (one multiply per sample).
__device__ __forceinline__ float freq_step(int i, int n, float dw) {
    return (float)((i <= (n - 1) / 2) ? i : i - n) * dw;
}

// 1 - exp(-t) for t >= 0: a Taylor series below 1/4 (truncation t^8 / 8! < 4e-10 relative), where
// the subtraction would cancel, and expf (about 1 ulp) above.
__device__ __forceinline__ float one_minus_exp_neg(float t) {
    const float ser =
        t * fmaf(-t * 0.5f,
                 fmaf(-t * (1.f / 3.f),
                      fmaf(-t * 0.25f, fmaf(-t * 0.2f, fmaf(-t * (1.f / 6.f), fmaf(-t * (1.f / 7.f), 1.f, 1.f), 1.f), 1.f),
                           1.f),
                      1.f),
                 1.f);
    return t < 0.25f ? ser : 1.f - expf(-t);
}

// MODE_FWD_REFLECT: line y of the reflect-padded real image (length wp) -> out[y][x].
// MODE_PLAIN:       in[line][:] -> out[line][:] (forward).
// MODE_INV_FIRST:   line kx of the transposed spectrum (length hp): loads conj(S G_k mul), the
//                   transform is along ky, optional B-spline prefilter along y -> out[kx][y].
// MODE_INV_SECOND:  line y (length wp): transform along kx, optional prefilter along x, scale,
//                   conj -> out[y][x].
enum { MODE_FWD_REFLECT = 0, MODE_PLAIN = 1, MODE_INV_FIRST = 2, MODE_INV_SECOND = 3 };

struct LineArgs {
    const float *img;  // MODE_FWD_REFLECT: the w x h image
    int w, h;
    int other;         // MODE_INV_FIRST: the other padded side (wp), for the x frequency
    int mul;
    int prefilter;
    double scale;
};

// Padded LDS line layout (lines.hpp) with SH = 4; the twiddle table exp(-2 pi i e / n) is staged
// in LDS behind the two line buffers.
// Three workgroups share a CU (LDS): the register budget is capped at three wavefronts per SIMD.
template <int MODE, int SH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_lines(const float2 *__restrict__ in, float2 *__restrict__ out, PlanArgs P,
                                               const float2 *__restrict__ tw_g, Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    const int n = P.n;
    const int ln = lp_len<SH>(n);
    float2 *a = lds, *b = lds + ln, *tw = lds + 2 * ln;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = tw_g[i];
    const int line = blockIdx.x;
    __syncthreads();
    float2 *r;
    if (MODE == MODE_FWD_REFLECT) {
        const float *row = A.img + (size_t)reflect_idx(line - PAD, A.h) * A.w;
        const int w = A.w;
        r = fft_line_from<SH>([&](int x) { return make_float2(row[reflect_idx(x - PAD, w)], 0.f); }, a, b, n, P.fac,
                              P.magic, P.nfac, tw, threadIdx.x, blockDim.x);
    } else if (MODE == MODE_INV_FIRST) {
        // G_k = (1 - exp(-|f|^2 s_hp^2 / 2)) exp(-|f - w0_k|^2 / (2 s_k^2)) (phaseflow.py:73-89)
        const float2 *src = in + (size_t)line * n;
        const float fx = freq(line, A.other);
        const float ax = fx - bank.w0x[k];
        const float w0y = bank.w0y[k], i2s = bank.inv2s2[k], shp = bank.s_hp2_half;
        const float dw = (float)(6.283185307179586 / (double)n);
        const float gx = ax * ax * i2s, hx = fx * fx * shp;
        const int mul = A.mul;
        r = fft_line_from<SH>(
            [&](int y) {
                const float fy = freq_step(y, n, dw);
                const float hp = one_minus_exp_neg(fmaf(fy * fy, shp, hx));
                const float ay = fy - w0y;
                const float G = hp * expf(-fmaf(ay * ay, i2s, gx));
                float2 v = cscale(src[y], G);
                if (mul == MUL_G_IFX) v = make_float2(-v.y * fx, v.x * fx);
                else if (mul == MUL_G_IFY) v = make_float2(-v.y * fy, v.x * fy);
                return cconj(v);
            },
            a, b, n, P.fac, P.magic, P.nfac, tw, threadIdx.x, blockDim.x);
    } else {
        const float2 *src = in + (size_t)line * n;
        r = fft_line_from<SH>([&](int x) { return src[x]; }, a, b, n, P.fac, P.magic, P.nfac, tw, threadIdx.x,
                              blockDim.x);
    }
    float2 *o = out + (size_t)line * n;
    if ((MODE == MODE_INV_FIRST || MODE == MODE_INV_SECOND) && A.prefilter) {
        float2 *sc = (r == a) ? b : a;
        prefilter_lines<float2, SH, Sgl2>(r, sc, r, n, 1, MODE == MODE_INV_SECOND ? A.scale : 1.0, threadIdx.x,
                                          blockDim.x);
        for (int x = threadIdx.x; x < n; x += blockDim.x)
            o[x] = MODE == MODE_INV_SECOND ? cconj(r[lp<SH>(x)]) : r[lp<SH>(x)];
    } else if (MODE == MODE_INV_SECOND) {
        const float s = (float)A.scale;
        for (int x = threadIdx.x; x < n; x += blockDim.x) o[x] = cconj(cscale(r[lp<SH>(x)], s));
    } else {
        for (int x = threadIdx.x; x < n; x += blockDim.x) o[x] = r[lp<SH>(x)];
    }
}

// out[c][r] = in[r][c] for an R x C array, 32 x 32 tiles through LDS.
__global__ void __launch_bounds__(256) k_transpose(const float2 *__restrict__ in, float2 *__restrict__ out, int R,
                                                   int C) {
    __shared__ float2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + tx;
        if (r < R && c < C) tile[j][tx] = in[(size_t)r * C + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + tx;
        if (r < R && c < C) out[(size_t)c * R + r] = tile[tx][j];
    }
}

std::mutex g_plan_mu;
std::map<int, FftPlan> g_plans;

}  // namespace

const FftPlan &fft_plan(int n, hipStream_t s) {
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
    // exp(-2 pi i e / n) from the double angle, rounded to float; ready before any other lane
    // can find the plan
    HIPCHECK(hipMalloc(&p.tw, n * sizeof(float2)));
    make_twiddles(p.tw, n, s);
    wait_stream(s);
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
// LDS of a line workgroup: two padded line buffers and the twiddle table (n <= 4096: at most
// 102 KiB, within the 160 KiB a CDNA4 workgroup may hold).
size_t line_lds_bytes(int n) { return (2 * (size_t)lp_len<4>(n) + (size_t)n) * sizeof(float2); }

template <int MODE>
void launch_lines(const float2 *in, float2 *out, int n, int lines, const Bank *bank, int k, const LineArgs &A,
                  hipStream_t s) {
    const FftPlan &P = fft_plan(n, s);
    const size_t lds = line_lds_bytes(n);
    if (lds > 65536) {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)k_lines<MODE, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
        });
    }
    Bank dummy{};
    hipLaunchKernelGGL((k_lines<MODE, 4>), dim3(lines), dim3(256), lds, s, in, out, args_of(P), P.tw,
                       bank ? *bank : dummy, k, A);
    LAUNCHCHECK();
}

void transpose(const float2 *in, float2 *out, int R, int C, hipStream_t s) {
    hipLaunchKernelGGL(k_transpose, dim3((C + 31) / 32, (R + 31) / 32), dim3(256), 0, s, in, out, R, C);
    LAUNCHCHECK();
}
}  // namespace

void fft_forward_padded(const float *img, int w, int h, float2 *spec, float2 *tmp, float2 *tmp2, hipStream_t s) {
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    LineArgs A{};
    A.img = img;
    A.w = w;
    A.h = h;
    launch_lines<MODE_FWD_REFLECT>(nullptr, tmp, wp, hp, nullptr, 0, A, s);  // [y][kx]
    transpose(tmp, tmp2, hp, wp, s);                                        // [kx][y]
    launch_lines<MODE_PLAIN>(tmp2, spec, hp, wp, nullptr, 0, A, s);         // [kx][ky]
}

void fft_inverse_filtered(const float2 *spec, int hp, int wp, const Bank &bank, int k, SpecMul mul, bool prefilter,
                          float2 *tmp, float2 *tmp2, float2 *out, hipStream_t s) {
    LineArgs A{};
    A.other = wp;
    A.mul = (int)mul;
    A.prefilter = prefilter ? 1 : 0;
    A.scale = 1.0 / ((double)hp * (double)wp);
    launch_lines<MODE_INV_FIRST>(spec, tmp, hp, wp, &bank, k, A, s);  // [kx][y]
    transpose(tmp, tmp2, wp, hp, s);                                 // [y][kx]
    launch_lines<MODE_INV_SECOND>(tmp2, out, wp, hp, nullptr, 0, A, s);
}

void fft_release() {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    for (auto &kv : g_plans) (void)hipFree(kv.second.tw);
    g_plans.clear();
}

}  // namespace ofx
