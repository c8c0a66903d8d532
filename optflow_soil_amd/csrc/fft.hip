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

// 1 - exp(-t) for t >= 0. Below 1/4, where the subtraction would cancel, the Taylor polynomial
// t (1 - t / 2 + t^2 / 6 - ... + t^6 / 5040) in Horner form (truncation t^8 / 8! < 4e-10
// relative); above, 1 - exp(-t) with the hardware base-2 exponential (__expf: a few 1e-7 relative
// error for the arguments of the filter bank, where the result is at least 0.22).
__device__ __forceinline__ float one_minus_exp_neg(float t) {
    float p = fmaf(t, 1.f / 5040.f, -1.f / 720.f);
    p = fmaf(t, p, 1.f / 120.f);
    p = fmaf(t, p, -1.f / 24.f);
    p = fmaf(t, p, 1.f / 6.f);
    p = fmaf(t, p, -0.5f);
    p = fmaf(t, p, 1.f);
    return t < 0.25f ? t * p : 1.f - __expf(-t);
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
    // a batch of planes (blockIdx.y): element offsets between consecutive input / output planes;
    // plane j of an inverse uses filter k + j
    size_t in_plane, out_plane;
};

// The transform of one line whose element e is load(e), by direct Stockham passes over LDS buffers
// a / b with the twiddle table tw; returns the buffer holding the result.
template <int SH>
struct DirectCore {
    static constexpr bool IN_PLACE = false;
    const PlanArgs &P;
    const float2 *tw;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *b, int tid, int nt) const {
        return fft_line_from<SH>(load, a, b, P.n, P.fac, P.magic, P.nfac, tw, tid, nt);
    }
};

// Chirp-z (Bluestein) form for lengths with a large prime factor, whose direct pass would cost
// n p operations per line: X_k = conj(b_k) sum_j (x_j conj(b_j)) b_{k-j} with b_j = exp(i pi j^2 / n)
// is a cyclic convolution of length m >= 2n - 1, done with two length-m transforms:
// A = FFT_m(x conj(b)), then conv = conj(FFT_m(conj(A C))) with C the length-m DFT of the wrapped
// chirp divided by m, so X_k = conj(b_k FFT_m(conj(A C))_k). The tables live in global memory
// (L2-resident); the LDS holds only the two line buffers of length m.
struct BlueArgs {
    PlanArgs M;            // the plan of length m
    const float2 *tw;      // exp(-2 pi i e / m)
    const float2 *chirp;   // b_j, j < n
    const float2 *spec;    // C / m
};

template <int SH>
struct BlueCore {
    static constexpr bool IN_PLACE = false;
    const BlueArgs &B;
    int n;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *b, int tid, int nt) const {
        const float2 *ch = B.chirp, *C = B.spec;
        const int n_ = n;
        float2 *r = fft_line_from<SH>(
            [&](int e) { return e < n_ ? cmul(load(e), cconj(ch[e])) : make_float2(0.f, 0.f); }, a, b, B.M.n,
            B.M.fac, B.M.magic, B.M.nfac, B.tw, tid, nt);
        float2 *o = (r == a) ? b : a;
        // the first pass reads r through the loader and writes o, so r is free for the later passes
        float2 *r2 = fft_line_from<SH>([&](int e) { return cconj(cmul(r[lp<SH>(e)], C[e])); }, r, o, B.M.n,
                                       B.M.fac, B.M.magic, B.M.nfac, B.tw, tid, nt);
        for (int k = tid; k < n_; k += nt) r2[lp<SH>(k)] = cconj(cmul(ch[k], r2[lp<SH>(k)]));
        __syncthreads();
        return r2;
    }
};

// One line per workgroup: load (with the mode's fused pre-processing), transform, post-process,
// store. The LDS buffers a / b hold at least n elements each (padded layout, lines.hpp).
template <int MODE, int SH, class Core>
__device__ __forceinline__ void line_body(const float2 *__restrict__ in, float2 *__restrict__ out, int n,
                                          const Bank &bank, int k, const LineArgs &A, float2 *a, float2 *b,
                                          const Core &core) {
    const int line = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    float2 *r;
    if (MODE == MODE_FWD_REFLECT) {
        const float *row = A.img + (size_t)reflect_idx(line - PAD, A.h) * A.w;
        const int w = A.w;
        r = core([&](int x) { return make_float2(row[reflect_idx(x - PAD, w)], 0.f); }, a, b, tid, nt);
    } else if (MODE == MODE_INV_FIRST) {
        // G_k = (1 - exp(-|f|^2 s_hp^2 / 2)) exp(-|f - w0_k|^2 / (2 s_k^2)) (phaseflow.py:73-89)
        const float2 *src = in + (size_t)line * n;
        const float fx = freq(line, A.other);
        const float ax = fx - bank.w0x[k];
        const float w0y = bank.w0y[k], i2s = bank.inv2s2[k], shp = bank.s_hp2_half;
        const float dw = (float)(6.283185307179586 / (double)n);
        const float gx = ax * ax * i2s, hx = fx * fx * shp;
        const int mul = A.mul;
        // the factor i fx or i fy by selects, not branches: a branch per element would split the
        // unrolled loads of a butterfly into blocks, each waiting for its own load
        r = core(
            [&](int y) {
                const float fy = freq_step(y, n, dw);
                const float hp = one_minus_exp_neg(fmaf(fy * fy, shp, hx));
                const float ay = fy - w0y;
                // the Gaussian factor by the hardware exponential: relative error a few 1e-7
                // where G matters (|argument| < 20), negligible in absolute terms beyond
                const float G = hp * __expf(-fmaf(ay * ay, i2s, gx));
                const float2 v = cscale(src[y], G);
                const float m = mul == MUL_G_IFX ? fx : fy;
                const float2 d = make_float2(-v.y * m, v.x * m);
                return cconj(mul == MUL_G ? v : d);
            },
            a, b, tid, nt);
    } else {
        const float2 *src = in + (size_t)line * n;
        r = core([&](int x) { return src[x]; }, a, b, tid, nt);
    }
    float2 *o = out + (size_t)line * n;
    if ((MODE == MODE_INV_FIRST || MODE == MODE_INV_SECOND) && A.prefilter) {
        const double scale = MODE == MODE_INV_SECOND ? A.scale : 1.0;
        if constexpr (Core::IN_PLACE) {
            prefilter_line_inplace<SH>(r, n, scale, tid, nt);
        } else {
            float2 *sc = (r == a) ? b : a;
            prefilter_lines<float2, SH, Sgl2>(r, sc, r, n, 1, scale, tid, nt);
        }
        for (int x = tid; x < n; x += nt) o[x] = MODE == MODE_INV_SECOND ? cconj(r[lp<SH>(x)]) : r[lp<SH>(x)];
    } else if (MODE == MODE_INV_SECOND) {
        const float s = (float)A.scale;
        for (int x = tid; x < n; x += nt) o[x] = cconj(cscale(r[lp<SH>(x)], s));
    } else {
        for (int x = tid; x < n; x += nt) o[x] = r[lp<SH>(x)];
    }
}

// Direct passes: the twiddle table is staged in LDS behind the two line buffers. Three workgroups
// share a CU (LDS): the register budget is capped at three wavefronts per SIMD.
template <int MODE, int SH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_lines(const float2 *__restrict__ in, float2 *__restrict__ out, PlanArgs P,
                                               const float2 *__restrict__ tw_g, Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    const int n = P.n;
    const int ln = lp_len<SH>(n);
    float2 *a = lds, *b = lds + ln;
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, n, bank, k + (int)blockIdx.y, A,
                        a, b, DirectCore<SH>{P, tw_g});
}

// The compile-time form of a line length 16 Rs... (lines.hpp fft_line_ct), 256 threads per line.
template <int N, int SH, int... Rs>
struct CtCore {
    static constexpr bool IN_PLACE = false;
    const float2 *T, *tw;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *b, int tid, int) const {
        return fft_line_ct<N, SH, 256, Rs...>(load, a, b, T, tw, tid);
    }
};

template <int MODE, int SH, int N, int... Rs>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_lines_ct(const float2 *__restrict__ in, float2 *__restrict__ out, const float2 *__restrict__ T,
                                                  const float2 *__restrict__ tw_g, Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    constexpr int ln = lp_len<SH>(N);
    float2 *a = lds, *b = lds + ln;
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, N, bank, k + (int)blockIdx.y, A,
                        a, b, CtCore<N, SH, Rs...>{T, tw_g});
}

// The compile-time form on one LDS buffer per line (lines.hpp fft_line_ct_inplace): 18 KB of LDS
// at 2112 instead of 36, so eight lines share a CU instead of four.
template <int N, int SH, int... Rs>
struct CtCore1 {
    static constexpr bool IN_PLACE = true;
    const float2 *T, *tw;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *, int tid, int) const {
        fft_line_ct_inplace<N, SH, 256, false, Rs...>(load, a, T, tw, tid);
        return a;
    }
};

template <int MODE, int SH, int N, int... Rs>
__global__ void __launch_bounds__(256) k_lines_ct1(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                   const float2 *__restrict__ T, const float2 *__restrict__ tw_g,
                                                   Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, N, bank, k + (int)blockIdx.y, A,
                        lds, nullptr, CtCore1<N, SH, Rs...>{T, tw_g});
}

// The second pass of an inverse straight from the first pass's layout [kx][y] (rows of `rows`
// samples), with no transpose between the passes. A workgroup of 4 x 256 threads takes the lines
// y0 .. y0 + 3: it loads them as 32-byte pieces of the rows (the 4 samples y0 .. y0 + 3 of one kx;
// a wavefront reads 16 such pieces), transposing through LDS, then each 256 threads transform one
// line in place. The 4 groups whose pieces share 128-byte rows segments (16 lines) run on one XCD
// one after another: slot s holds group (s mod 8) (slots / 8) + s / 8 (workgroups are dealt to the
// 8 XCDs in turn), so their reads meet in that XCD's L2.
// MODE_PLAIN runs the forward transform's second pass the same way (lines kx from [y][kx]).
constexpr int TL = 4;
template <int MODE, int SH, int N, int... Rs>
__global__ void __launch_bounds__(1024) k_lines_ct1_t(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                      const float2 *__restrict__ T, const float2 *__restrict__ tw_g,
                                                      int rows, int groups, LineArgs A) {
    extern __shared__ float2 lds[];
    const int slot = blockIdx.x;
    const int g = (slot & 7) * ((int)gridDim.x >> 3) + (slot >> 3);
    if (g >= groups) return;  // a padding slot: the whole workgroup leaves before any barrier
    constexpr int LN = lp_len<SH>(N), LOADS = (N * TL + 1023) / 1024;
    const int y0 = g * TL, j = threadIdx.x >> 8, tid = threadIdx.x & 255;
    in += blockIdx.y * A.in_plane;
    out += blockIdx.y * A.out_plane;
    float2 v[LOADS];
#pragma unroll
    for (int t = 0; t < LOADS; t++) {
        const int e = t * 1024 + threadIdx.x, kx = e / TL, y = y0 + e % TL;
        v[t] = (e < N * TL && y < rows) ? in[(size_t)kx * rows + y] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int t = 0; t < LOADS; t++) {
        const int e = t * 1024 + threadIdx.x;
        if (e < N * TL) lds[(e % TL) * LN + lp<SH>(e / TL)] = v[t];
    }
    __syncthreads();
    float2 *buf = lds + j * LN;
    fft_line_ct_inplace<N, SH, 256, true, Rs...>([&](int x) { return buf[lp<SH>(x)]; }, buf, T, tw_g, tid);
    const int y = y0 + j;
    if (MODE == MODE_INV_SECOND && A.prefilter) prefilter_line_inplace<SH>(buf, N, A.scale, tid, 256);
    if (y >= rows) return;
    float2 *o = out + (size_t)y * N;
    if (MODE == MODE_PLAIN) {
        for (int x = tid; x < N; x += 256) o[x] = buf[lp<SH>(x)];
    } else if (A.prefilter) {
        for (int x = tid; x < N; x += 256) o[x] = cconj(buf[lp<SH>(x)]);
    } else {
        const float sc = (float)A.scale;
        for (int x = tid; x < N; x += 256) o[x] = cconj(cscale(buf[lp<SH>(x)], sc));
    }
}

// T[k (R - 1) + q - 1] = tw[q k step] for k < p, 0 < q < R: a pass table of the compile-time form
__global__ void k_ct_twiddles(float2 *__restrict__ T, const float2 *__restrict__ tw, int p, int R, int step) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p * (R - 1)) return;
    const int k = e / (R - 1), q = e - k * (R - 1) + 1;
    T[e] = tw[q * k * step];
}

// The lengths with a compile-time form: the plan's factors exactly (16, r[0], r[1], ...). 2112,
// 1088, 576 and 832 are the padded sides of 2048, 1024, 512 and 768 px frames; 1216 that of a
// 1024 px tile with its margins; 1600 that of 1536 px; 8192 and 6144 the chirp-z lengths of the
// 4000 x 3000 frame's sides (4064 and 3064 padded).
struct CtForm {
    int n, nr, r[3];
};
constexpr CtForm CT_FORMS[] = {{2112, 2, {12, 11}}, {1088, 2, {4, 17}}, {576, 2, {12, 3}},  {832, 2, {4, 13}},
                               {1216, 2, {4, 19}},  {1600, 3, {4, 5, 5}}, {8192, 3, {8, 8, 8}}, {6144, 3, {12, 8, 4}}};

// The chirp-z form: the two line buffers of length m (m <= 8192: at most 136 KiB) leave room for
// one workgroup per CU, so a workgroup of 1024 threads keeps 16 wavefronts on the CU.
template <int MODE, int SH>
__global__ void __launch_bounds__(1024) k_lines_blue(const float2 *__restrict__ in, float2 *__restrict__ out, int n,
                                                     BlueArgs B, Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    const int lm = lp_len<SH>(B.M.n);
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, n, bank, k + (int)blockIdx.y, A,
                        lds, lds + lm, BlueCore<SH>{B, n});
}

// The chirp-z form with its length-m transforms in the compile-time form (m = 16 Rs...).
template <int M, int SH, int... Rs>
struct BlueCtCore {
    static constexpr bool IN_PLACE = false;
    const float2 *T, *tw, *ch, *C;
    int n;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *b, int tid, int) const {
        const int n_ = n;
        const float2 *ch_ = ch, *C_ = C;
        float2 *r = fft_line_ct<M, SH, 1024, Rs...>(
            [&](int e) { return e < n_ ? cmul(load(e), cconj(ch_[e])) : make_float2(0.f, 0.f); }, a, b, T, tw, tid);
        float2 *o = (r == a) ? b : a;
        // the first pass reads r through the loader and writes o, so r is free for the later passes
        float2 *r2 = fft_line_ct<M, SH, 1024, Rs...>([&](int e) { return cconj(cmul(r[lp<SH>(e)], C_[e])); }, o, r, T,
                                                     tw, tid);
        for (int k = tid; k < n_; k += 1024) r2[lp<SH>(k)] = cconj(cmul(ch_[k], r2[lp<SH>(k)]));
        __syncthreads();
        return r2;
    }
};

template <int MODE, int SH, int M, int... Rs>
__global__ void __launch_bounds__(1024) k_lines_blue_ct(const float2 *__restrict__ in, float2 *__restrict__ out, int n,
                                                        const float2 *__restrict__ T, const float2 *__restrict__ tw,
                                                        const float2 *__restrict__ ch, const float2 *__restrict__ C,
                                                        Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    constexpr int lm = lp_len<SH>(M);
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, n, bank, k + (int)blockIdx.y, A,
                        lds, lds + lm, BlueCtCore<M, SH, Rs...>{T, tw, ch, C, n});
}

// The chirp-z form on one LDS buffer of length m (lines.hpp fft_line_ct_inplace): 70 KB at 8192
// instead of 139, so two 1024-thread workgroups share a CU instead of one.
template <int M, int SH, int... Rs>
struct BlueCtCore1 {
    static constexpr bool IN_PLACE = true;
    const float2 *T, *tw, *ch, *C;
    int n;
    template <class L>
    __device__ float2 *operator()(L load, float2 *a, float2 *, int tid, int) const {
        const int n_ = n;
        const float2 *ch_ = ch, *C_ = C;
        fft_line_ct_inplace<M, SH, 1024, false, Rs...>(
            [&](int e) { return e < n_ ? cmul(load(e), cconj(ch_[e])) : make_float2(0.f, 0.f); }, a, T, tw, tid);
        fft_line_ct_inplace<M, SH, 1024, true, Rs...>([&](int e) { return cconj(cmul(a[lp<SH>(e)], C_[e])); }, a, T,
                                                      tw, tid);
        for (int k = tid; k < n_; k += 1024) a[lp<SH>(k)] = cconj(cmul(ch_[k], a[lp<SH>(k)]));
        __syncthreads();
        return a;
    }
};

template <int MODE, int SH, int M, int... Rs>
__global__ void __launch_bounds__(1024) k_lines_blue_ct1(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                         int n, const float2 *__restrict__ T,
                                                         const float2 *__restrict__ tw, const float2 *__restrict__ ch,
                                                         const float2 *__restrict__ C, Bank bank, int k, LineArgs A) {
    extern __shared__ float2 lds[];
    line_body<MODE, SH>(in + blockIdx.y * A.in_plane, out + blockIdx.y * A.out_plane, n, bank, k + (int)blockIdx.y, A,
                        lds, nullptr, BlueCtCore1<M, SH, Rs...>{T, tw, ch, C, n});
}

// b_j = exp(i pi (j^2 mod 2n) / n) from the double angle (the product j^2 is exact in 64 bits)
__global__ void k_chirp(float2 *__restrict__ b, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long long t = ((long long)j * j) % (2LL * n);
    const double ph = M_PI * (double)t / (double)n;
    b[j] = make_float2((float)cos(ph), (float)sin(ph));
}

// C_k = (1 / m) sum_j c_j exp(-2 pi i j k / m) for the wrapped chirp c (c_j = b_j for j < n,
// c_{m-j} = b_j for 0 < j < n, 0 between), summed in double by direct evaluation (once per length)
__global__ void k_chirp_spectrum(float2 *__restrict__ C, int n, int m) {
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    if (kk >= m) return;
    double sr = 0, si = 0;
    for (int j = 0; j < m; j++) {
        const int jj = j < n ? j : (j > m - n ? m - j : -1);
        if (jj < 0) continue;
        const long long t = ((long long)jj * jj) % (2LL * n);
        const long long e = ((long long)j * kk) % m;
        const double ph = M_PI * (double)t / (double)n - 2.0 * M_PI * (double)e / (double)m;
        sr += cos(ph);
        si += sin(ph);
    }
    C[kk] = make_float2((float)(sr / m), (float)(si / m));
}

// out[c][r] = in[r][c] for an R x C array, 64 x 64 tiles through LDS; blockIdx.z: the plane of a
// batch of such arrays stored one after the other. A wavefront reads 512 contiguous bytes of a row
// and writes 512 of a column; each thread issues its 16 loads before the first LDS write, so a CU
// (four such workgroups, 33 KB of LDS each) keeps about 128 KB of reads in flight.
constexpr int TR_T = 64;
__global__ void __launch_bounds__(256) k_transpose(const float2 *__restrict__ in, float2 *__restrict__ out, int R,
                                                   int C) {
    __shared__ float2 tile[TR_T][TR_T + 1];
    const int c0 = blockIdx.x * TR_T, r0 = blockIdx.y * TR_T;
    in += (size_t)blockIdx.z * R * C;
    out += (size_t)blockIdx.z * R * C;
    const int tx = threadIdx.x & (TR_T - 1), ty = threadIdx.x >> 6;  // 64 x 4
    float2 v[TR_T / 4];
#pragma unroll
    for (int k = 0; k < TR_T / 4; k++) {
        const int r = r0 + ty + 4 * k, c = c0 + tx;
        v[k] = (r < R && c < C) ? in[(size_t)r * C + c] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < TR_T / 4; k++) tile[ty + 4 * k][tx] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TR_T / 4; k++) {
        const int c = c0 + ty + 4 * k, r = r0 + tx;
        if (r < R && c < C) out[(size_t)c * R + r] = tile[tx][ty + 4 * k];
    }
}

std::mutex g_plan_mu;
std::map<int, FftPlan> g_plans;

}  // namespace

namespace {
// Factors and twiddles of a length-n plan; the caller holds g_plan_mu.
const FftPlan &plan_locked(int n, int max_n, hipStream_t s) {
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    if (n < 1 || n > max_n)
        throw Error(OFX_ETOOLARGE, "FFT length " + std::to_string(n) + " exceeds the " + std::to_string(max_n) +
                                       " samples an LDS line holds");
    FftPlan p;
    p.n = n;
    int m = n, big = 1;
    auto push = [&](int r) {
        if (p.nfac >= MAX_FACTORS) throw Error(OFX_EUNSUPPORTED, "FFT length has too many factors");
        p.fac[p.nfac++] = r;
        m /= r;
        big = std::max(big, r);
    };
    // a radix-16 first pass (it needs no twiddles) and a radix-12 pass: fewer passes (2112 = 16 12 11
    // in three instead of 8 8 3 11 in four), so fewer LDS round trips and barriers per line
    if (m % 16 == 0) push(16);
    if (m % 12 == 0) push(12);
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
    for (const CtForm &f : CT_FORMS) {
        bool match = f.n == n && p.nfac == f.nr + 1 && p.fac[0] == 16;
        for (int j = 0; match && j < f.nr; j++) match = p.fac[j + 1] == f.r[j];
        if (!match) continue;
        size_t total = 0;
        for (int j = 0, pp = 16; j < f.nr; pp *= f.r[j], j++) total += (size_t)pp * (f.r[j] - 1);
        HIPCHECK(hipMalloc(&p.ct_tw, total * sizeof(float2)));
        size_t off = 0;
        for (int j = 0, pp = 16; j < f.nr; pp *= f.r[j], j++) {
            const int cnt = pp * (f.r[j] - 1);
            hipLaunchKernelGGL(k_ct_twiddles, dim3((cnt + 255) / 256), dim3(256), 0, s, p.ct_tw + off, p.tw, pp,
                               f.r[j], n / (pp * f.r[j]));
            off += cnt;
        }
        LAUNCHCHECK();
    }
    if (big >= BLUE_MIN_PRIME) {
        // the smallest 2^a 3^b >= 2n - 1 (at most 8192 for n <= 4096)
        int best = 0;
        for (long long p2 = 1; p2 <= 16384; p2 *= 2)
            for (long long v = p2; v <= 16384; v *= 3)
                if (v >= 2LL * n - 1 && (best == 0 || v < best)) best = (int)v;
        plan_locked(best, 2 * 4096, s);
        p.blue_m = best;
        HIPCHECK(hipMalloc(&p.chirp, n * sizeof(float2)));
        HIPCHECK(hipMalloc(&p.chirp_spec, best * sizeof(float2)));
        hipLaunchKernelGGL(k_chirp, dim3((n + 255) / 256), dim3(256), 0, s, p.chirp, n);
        hipLaunchKernelGGL(k_chirp_spectrum, dim3((best + 63) / 64), dim3(64), 0, s, p.chirp_spec, n, best);
        LAUNCHCHECK();
    }
    wait_stream(s);
    return g_plans[n] = p;
}
}  // namespace

const FftPlan &fft_plan(int n, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    return plan_locked(n, 4096, s);
}

namespace {
// The plan of a chirp-z convolution length (up to 8192; created with the plan that uses it)
const FftPlan &fft_plan_m(int m, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    return plan_locked(m, 2 * 4096, s);
}
}  // namespace

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
size_t line_lds_bytes(int n) { return 2 * (size_t)lp_len<4>(n) * sizeof(float2); }

template <int MODE>
void launch_lines(const float2 *in, float2 *out, int n, int lines, int planes, const Bank *bank, int k,
                  const LineArgs &A, hipStream_t s) {
    const FftPlan &P = fft_plan(n, s);
    Bank dummy{};
    if (P.blue_m) {
        const FftPlan &M = fft_plan_m(P.blue_m, s);
        const size_t lds = 2 * (size_t)lp_len<4>(P.blue_m) * sizeof(float2);
        if (M.ct_tw) {
            const Bank bk = bank ? *bank : dummy;
            const size_t lds1 = (size_t)lp_len<4>(P.blue_m) * sizeof(float2);
#define OFX_BLUE_CT_LAUNCH(M_, ...)                                                                                 \
    if (P.blue_m == M_) {                                                                                           \
        static std::once_flag once_ct;                                                                              \
        std::call_once(once_ct, [] {                                                                                \
            (void)hipFuncSetAttribute((const void *)k_lines_blue_ct1<MODE, 4, M_, __VA_ARGS__>,                     \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);                      \
        });                                                                                                         \
        hipLaunchKernelGGL((k_lines_blue_ct1<MODE, 4, M_, __VA_ARGS__>), dim3(lines, planes), dim3(1024), lds1, s,   \
                           in, out, n, M.ct_tw, M.tw, P.chirp, P.chirp_spec, bk, k, A);                              \
        LAUNCHCHECK();                                                                                              \
        return;                                                                                                     \
    }
            OFX_BLUE_CT_LAUNCH(8192, 8, 8, 8)
            OFX_BLUE_CT_LAUNCH(6144, 12, 8, 4)
#undef OFX_BLUE_CT_LAUNCH
        }
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)k_lines_blue<MODE, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
        });
        const BlueArgs B{args_of(M), M.tw, P.chirp, P.chirp_spec};
        hipLaunchKernelGGL((k_lines_blue<MODE, 4>), dim3(lines, planes), dim3(1024), lds, s, in, out, n, B,
                           bank ? *bank : dummy, k, A);
        LAUNCHCHECK();
        return;
    }
    const size_t lds = line_lds_bytes(n);
    if (P.ct_tw) {
        const Bank bk = bank ? *bank : dummy;
#define OFX_CT_LAUNCH(N_, ...)                                                                                      \
    if (n == N_) {                                                                                                  \
        hipLaunchKernelGGL((k_lines_ct1<MODE, 4, N_, __VA_ARGS__>), dim3(lines, planes), dim3(256),                 \
                           lp_len<4>(N_) * sizeof(float2), s, in, out, P.ct_tw, P.tw, bk, k, A);                    \
        LAUNCHCHECK();                                                                                              \
        return;                                                                                                     \
    }
        OFX_CT_LAUNCH(2112, 12, 11)
        OFX_CT_LAUNCH(1088, 4, 17)
        OFX_CT_LAUNCH(576, 12, 3)
        OFX_CT_LAUNCH(832, 4, 13)
        OFX_CT_LAUNCH(1216, 4, 19)
        OFX_CT_LAUNCH(1600, 4, 5, 5)
#undef OFX_CT_LAUNCH
    }
    if (lds > 65536) {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)k_lines<MODE, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
        });
    }
    hipLaunchKernelGGL((k_lines<MODE, 4>), dim3(lines, planes), dim3(256), lds, s, in, out, args_of(P), P.tw,
                       bank ? *bank : dummy, k, A);
    LAUNCHCHECK();
}

void transpose(const float2 *in, float2 *out, int R, int C, int planes, hipStream_t s) {
    hipLaunchKernelGGL(k_transpose, dim3((C + TR_T - 1) / TR_T, (R + TR_T - 1) / TR_T, planes), dim3(256), 0, s, in,
                       out, R, C);
    LAUNCHCHECK();
}
}  // namespace

namespace {
// A second pass (lines of length n, `rows` of them) straight from the first pass's planes with
// k_lines_ct1_t, when n has a compile-time form; false otherwise (the caller transposes).
template <int MODE>
bool second_pass_from_columns(const float2 *in, float2 *out, int n, int rows, int planes, const LineArgs &A,
                              hipStream_t s) {
    const FftPlan &P = fft_plan(n, s);
    if (P.blue_m || !P.ct_tw) return false;
    const int groups = (rows + TL - 1) / TL, slots = 8 * ((groups + 7) / 8);
#define OFX_CT_T_LAUNCH(N_, ...)                                                                                    \
    if (n == N_) {                                                                                                  \
        static std::once_flag once_t;                                                                               \
        std::call_once(once_t, [] {                                                                                 \
            (void)hipFuncSetAttribute((const void *)k_lines_ct1_t<MODE, 4, N_, __VA_ARGS__>,                        \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);                      \
        });                                                                                                         \
        hipLaunchKernelGGL((k_lines_ct1_t<MODE, 4, N_, __VA_ARGS__>), dim3(slots, planes), dim3(1024),              \
                           TL * lp_len<4>(N_) * sizeof(float2), s, in, out, P.ct_tw, P.tw, rows, groups, A);        \
        LAUNCHCHECK();                                                                                              \
        return true;                                                                                                \
    }
    OFX_CT_T_LAUNCH(2112, 12, 11)
    OFX_CT_T_LAUNCH(576, 12, 3)
    OFX_CT_T_LAUNCH(832, 4, 13)
    OFX_CT_T_LAUNCH(1600, 4, 5, 5)
    // 1088 = 16 4 17 and 1216 = 16 4 19 keep the transpose: their radix-17 and -19 butterflies need
    // 78 VGPRs here, too many for two 1024-thread workgroups per CU
#undef OFX_CT_T_LAUNCH
    return false;
}
}  // namespace

void fft_forward_padded(const float *img, int w, int h, float2 *spec, float2 *tmp, float2 *tmp2, hipStream_t s) {
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    LineArgs A{};
    A.img = img;
    A.w = w;
    A.h = h;
    launch_lines<MODE_FWD_REFLECT>(nullptr, tmp, wp, hp, 1, nullptr, 0, A, s);  // [y][kx]
    A.in_plane = A.out_plane = 0;
    if (second_pass_from_columns<MODE_PLAIN>(tmp, spec, hp, wp, 1, A, s)) return;  // [kx][ky]
    transpose(tmp, tmp2, hp, wp, 1, s);                                         // [kx][y]
    launch_lines<MODE_PLAIN>(tmp2, spec, hp, wp, 1, nullptr, 0, A, s);          // [kx][ky]
}

void fft_inverse_filtered(const float2 *spec, int hp, int wp, const Bank &bank, int k0, int K, SpecMul mul,
                          bool prefilter, float2 *tmp, float2 *tmp2, float2 *out, hipStream_t s) {
    const size_t np = (size_t)hp * wp;
    LineArgs A{};
    A.other = wp;
    A.mul = (int)mul;
    A.prefilter = prefilter ? 1 : 0;
    A.scale = 1.0 / ((double)hp * (double)wp);
    A.in_plane = 0;  // every filter reads the one spectrum
    A.out_plane = np;
    launch_lines<MODE_INV_FIRST>(spec, tmp, hp, wp, K, &bank, k0, A, s);  // [kx][y] per filter
    A.in_plane = np;
    if (second_pass_from_columns<MODE_INV_SECOND>(tmp, out, wp, hp, K, A, s)) return;
    transpose(tmp, tmp2, wp, hp, K, s);  // [y][kx]
    launch_lines<MODE_INV_SECOND>(tmp2, out, wp, hp, K, nullptr, 0, A, s);
}

void fft_release() {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    for (auto &kv : g_plans) {
        (void)hipFree(kv.second.tw);
        (void)hipFree(kv.second.chirp);
        (void)hipFree(kv.second.chirp_spec);
        (void)hipFree(kv.second.ct_tw);
    }
    g_plans.clear();
}

}  // namespace ofx
