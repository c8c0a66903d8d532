// seq_mode "hybrid" on the GPU: the per-pixel state machine of the direct-first sequence with the
// gated adjacent-frame fallback (rust/crates/optflow-core/src/track/hybrid.rs:262-540,
// python/optflow_ref/track.py:383-505) and its local affine gate (hybrid.rs:147-260,
// track.py:331-380). The legs themselves are ordinary pairs; the host steps the frames.
#include <cmath>

#include "fine.hpp"

namespace ofx {

namespace {

dim3 grid2(int w, int h) { return dim3((w + 255) / 256, h); }
size_t nblk(size_t n) { return (n + 255) / 256; }

__device__ __forceinline__ bool in_roi(const uint8_t *mask, size_t i) { return !mask || mask[i]; }

// Counters aggregated per wavefront (one atomic per wavefront). Every lane of the wavefront must
// reach the call: the kernels below guard their memory accesses instead of returning early.
__device__ __forceinline__ void wave_count(int *c, bool pred) {
    const unsigned long long b = __ballot(pred);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(c, (int)__popcll(b));
}
__device__ __forceinline__ void wave_max(int *c, int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMax(c, v);
}

// A reference move: c0K is replaced (not min-accumulated) by the confidence of the leg old
// reference -> k - 1 where the pixel was direct-good at k - 1, and 0 where it was carried.
__global__ void k_hyb_refmove(const uint8_t *__restrict__ good_prev, const double *__restrict__ cp,
                              double *__restrict__ c0r, uint8_t *__restrict__ anch, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    c0r[i] = good_prev[i] ? cp[i] : 0.0;
    anch[i] = good_prev[i];
}

// link, direct_ok, good, re-acquisition candidates and the pixels that need the adjacent leg
__global__ void k_hyb_classify(const double *__restrict__ cx, const double *__restrict__ cy,
                               const double *__restrict__ leg_c, const double *__restrict__ c0r,
                               const uint8_t *__restrict__ anch, const uint8_t *__restrict__ mask,
                               const double *__restrict__ upx, const double *__restrict__ upy,
                               const int *__restrict__ nfb, double conf_min, int fb_max, int reacq_ok,
                               double *__restrict__ link, uint8_t *__restrict__ flags, int *__restrict__ counters,
                               size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    if (i < n) {
        const double l = anch[i] ? fmin(c0r[i], leg_c[i]) : leg_c[i];
        link[i] = l;
        const bool roi = in_roi(mask, i);
        const bool dok = isfinite(cx[i]) && isfinite(cy[i]) && l >= conf_min;
        const bool good = dok && anch[i];
        const bool cand_re = dok && !anch[i] && roi && reacq_ok;
        need = !good && roi && isfinite(upx[i]) && isfinite(upy[i]) && nfb[i] < fb_max;
        flags[i] = (uint8_t)((dok ? HYB_DIRECT_OK : 0) | (good ? HYB_GOOD : 0) | (cand_re ? HYB_CAND_RE : 0) |
                             (need ? HYB_NEED : 0));
    }
    wave_count(&counters[HYB_ANY_NEED], need);
}

// the carried value u_{k-1} + a(X + u_{k-1}); where the composition leaves the image the carried
// value stays
__global__ void k_hyb_carried(const double *__restrict__ cax, const double *__restrict__ cay,
                              const double *__restrict__ upx, const double *__restrict__ upy,
                              double *__restrict__ ax, double *__restrict__ ay, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool fin = isfinite(cax[i]) && isfinite(cay[i]);
    ax[i] = fin ? cax[i] : upx[i];
    ay[i] = fin ? cay[i] : upy[i];
}

// a fallback value exists; re-acquisition candidates without one
__global__ void k_hyb_has_fb(uint8_t *__restrict__ flags, const double *__restrict__ ca, double conf_min,
                             const uint8_t *__restrict__ mask, int *__restrict__ counters, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool count = false;
    if (i < n) {
        uint8_t f = flags[i];
        const bool has_fb = (f & HYB_NEED) && ca[i] >= conf_min;
        const bool re_nofb = (f & HYB_CAND_RE) && !has_fb;
        f |= (has_fb ? HYB_HAS_FB : 0) | (re_nofb ? HYB_RE_NOFB : 0);
        flags[i] = f;
        count = re_nofb && in_roi(mask, i);
    }
    wave_count(&counters[HYB_ANY_RE_NOFB], count);
}

// re-acquisition (gate0: the local-model test of the candidates without a fallback value, or
// NULL when the gate is off or was not needed), the fallback counter and the fallback candidates
__global__ void k_hyb_reacquire(uint8_t *__restrict__ flags, const uint8_t *__restrict__ gate0,
                                const double *__restrict__ cx, const double *__restrict__ cy,
                                const double *__restrict__ ax, const double *__restrict__ ay,
                                const double *__restrict__ leg_c, const double *__restrict__ ca, double fb_tol,
                                double conf_min, int fb_max, double *__restrict__ c0r, uint8_t *__restrict__ anch,
                                int *__restrict__ nfb, const uint8_t *__restrict__ mask, int *__restrict__ counters,
                                size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool count = false;
    if (i < n) {
        uint8_t f = flags[i];
        bool re_nofb = (f & HYB_RE_NOFB) != 0;
        if (gate0) re_nofb = re_nofb && gate0[i];
        const bool agree = hypot(cx[i] - ax[i], cy[i] - ay[i]) <= fb_tol;
        const bool re = ((f & HYB_CAND_RE) && (f & HYB_HAS_FB) && agree) || re_nofb;
        bool good = (f & HYB_GOOD) != 0;
        if (re) {
            c0r[i] = leg_c[i];
            anch[i] = 1;
            good = true;
        }
        const long cap = (long)fb_max + 1;
        const int nf = good ? 0 : (int)min((long)nfb[i] + 1, cap);
        nfb[i] = nf;
        const bool cand = !good && (f & HYB_NEED) && nf <= fb_max && ca[i] >= conf_min;
        f = (uint8_t)((f & ~(HYB_GOOD | HYB_RE_NOFB)) | (good ? HYB_GOOD : 0) | (re_nofb ? HYB_RE_NOFB : 0) |
                      (re ? HYB_RE : 0) | (cand ? HYB_FB_CAND : 0));
        flags[i] = f;
        count = cand && in_roi(mask, i);
    }
    wave_count(&counters[HYB_ANY_CAND], count);
}

// outputs of the frame: u (direct or carried), conf, the low mask of finish (neither direct-good
// nor an accepted fallback), fb, and the shares of the timing
__global__ void k_hyb_output(const uint8_t *__restrict__ flags, const uint8_t *__restrict__ gate,
                             const double *__restrict__ cx, const double *__restrict__ cy,
                             const double *__restrict__ ax, const double *__restrict__ ay,
                             const double *__restrict__ link, const double *__restrict__ ca,
                             const int *__restrict__ nfb, int fb_max, const uint8_t *__restrict__ mask,
                             double *__restrict__ ux, double *__restrict__ uy, float *__restrict__ conf,
                             uint8_t *__restrict__ low, int32_t *__restrict__ fb, int *__restrict__ counters,
                             size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool roi = false, fb_ok = false, rej = false, re = false;
    int nf = 0;
    if (i < n) {
        const uint8_t f = flags[i];
        const bool good = (f & HYB_GOOD) != 0, cand = (f & HYB_FB_CAND) != 0;
        fb_ok = cand && (!gate || gate[i]);
        rej = cand && !fb_ok;
        re = (f & HYB_RE) != 0;
        nf = nfb[i];
        ux[i] = good ? cx[i] : ax[i];
        uy[i] = good ? cy[i] : ay[i];
        const double cap = (double)fb_max + 1.0;
        conf[i] = good ? (float)link[i] : (float)(ca[i] * fmax(0.0, 1.0 - (double)nf / cap));
        low[i] = !(good || fb_ok);
        if (fb) fb[i] = nf;
        roi = in_roi(mask, i);
    }
    wave_max(&counters[HYB_NFB_MAX], nf);
    wave_count(&counters[HYB_N_FB_OK], roi && fb_ok);
    wave_count(&counters[HYB_N_REJECTED], roi && rej);
    wave_count(&counters[HYB_N_RE], roi && re);
    wave_count(&counters[HYB_N_ROI], roi);
}

// ---------------------------------------------------------------- local affine gate
// Row pass: the 12 window sums along x (half-width r, zero outside) of
// wt, wt x, wt y, wt x^2, wt x y, wt y^2, wt u0, wt x u0, wt y u0, wt v0, wt x v0, wt y v0
// with wt = sel and u0 = the value on the support (0 off it).
__global__ void k_gate_rows(const uint8_t *__restrict__ sel, const uint8_t *__restrict__ support,
                            const double *__restrict__ ux, const double *__restrict__ uy, int w, int h, int r,
                            double *__restrict__ rows) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h, base = (size_t)y * w;
    const double Y = (double)y;
    double s[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int xx = max(0, x - r); xx <= min(w - 1, x + r); xx++) {
        const size_t j = base + xx;
        if (!sel[j]) continue;
        const double X = (double)xx;
        const double u = support[j] ? ux[j] : 0.0, v = support[j] ? uy[j] : 0.0;
        s[0] += 1.0;
        s[1] += X;
        s[2] += Y;
        s[3] += X * X;
        s[4] += X * Y;
        s[5] += Y * Y;
        s[6] += u;
        s[7] += X * u;
        s[8] += Y * u;
        s[9] += v;
        s[10] += X * v;
        s[11] += Y * v;
    }
#pragma unroll
    for (int q = 0; q < 12; q++) rows[q * n + base + x] = s[q];
}

// m c = b (two right-hand sides) by Gaussian elimination with partial pivoting, the last of equal
// maxima as pivot (hybrid.rs:114-145); false if singular.
__device__ bool solve3(double m[3][3], double b[3][2], double xo[3][2]) {
    double a[3][5];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) a[i][j] = m[i][j];
        a[i][3] = b[i][0];
        a[i][4] = b[i][1];
    }
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int i = c + 1; i < 3; i++)
            if (!(fabs(a[i][c]) < fabs(a[p][c]))) p = i;
        if (a[p][c] == 0.0 || !isfinite(a[p][c])) return false;
        if (p != c)
            for (int j = 0; j < 5; j++) {
                const double t = a[c][j];
                a[c][j] = a[p][j];
                a[p][j] = t;
            }
        for (int i = c + 1; i < 3; i++) {
            const double f = a[i][c] / a[c][c];
            for (int j = c; j < 5; j++) a[i][j] -= f * a[c][j];
        }
    }
    for (int k = 0; k < 2; k++)
        for (int i = 2; i >= 0; i--) {
            double s = a[i][3 + k];
            for (int j = i + 1; j < 3; j++) s -= a[i][j] * xo[j][k];
            xo[i][k] = s / a[i][i];
        }
    return true;
}

// Column pass and fit: one thread per (column, segment of GATE_SEG rows) keeps the 12 column
// window sums as running sums (added below, dropped above, as the reference's box sum does) and
// fits the affine model at every pixel of its segment.
constexpr int GATE_SEG = 64;

__global__ void k_gate_cols_fit(const double *__restrict__ rows, int w, int h, int r, double ridge,
                                double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ s1o,
                                uint8_t *__restrict__ ok_fit) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, seg = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    const int y0 = seg * GATE_SEG, y1 = min(y0 + GATE_SEG, h);
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0.0;
    for (int yy = max(0, y0 - r); yy <= min(h - 1, y0 + r); yy++)
#pragma unroll
        for (int q = 0; q < 12; q++) acc[q] += rows[q * n + (size_t)yy * w + x];
    const double X = (double)x;
    for (int y = y0; y < y1; y++) {
        const size_t i = (size_t)y * w + x;
        const double Y = (double)y;
        const double S1 = acc[0];
        s1o[i] = S1;
        if (S1 < 3.0) {
            ax[i] = ay[i] = NAN;
            ok_fit[i] = 0;
        } else {
            const double sxc = acc[1] - X * S1, syc = acc[2] - Y * S1;
            const double sxxc = acc[3] - 2.0 * X * acc[1] + X * X * S1;
            const double sxyc = acc[4] - X * acc[2] - Y * acc[1] + X * Y * S1;
            const double syyc = acc[5] - 2.0 * Y * acc[2] + Y * Y * S1;
            double m[3][3] = {{S1, sxc, syc}, {sxc, sxxc + ridge * S1, sxyc}, {syc, sxyc, syyc + ridge * S1}};
            double b[3][2], c[3][2];
            for (int k = 0; k < 2; k++) {
                const double su = acc[6 + 3 * k];
                b[0][k] = su;
                b[1][k] = acc[7 + 3 * k] - X * su;
                b[2][k] = acc[8 + 3 * k] - Y * su;
            }
            if (solve3(m, b, c)) {
                ax[i] = c[0][0];
                ay[i] = c[0][1];
            } else {
                ax[i] = ay[i] = NAN;
            }
            ok_fit[i] = 1;
        }
        // slide the window: add row y + r + 1, drop row y - r
        const int add = y + r + 1, drop = y - r;
        if (add < h)
#pragma unroll
            for (int q = 0; q < 12; q++) acc[q] += rows[q * n + (size_t)add * w + x];
        if (drop >= 0)
#pragma unroll
            for (int q = 0; q < 12; q++) acc[q] -= rows[q * n + (size_t)drop * w + x];
    }
}

// support pixels whose residual against their own local model exceeds 3 tol are dropped
__global__ void k_gate_reselect(const uint8_t *__restrict__ support, const double *__restrict__ ux,
                                const double *__restrict__ uy, const double *__restrict__ ax,
                                const double *__restrict__ ay, double tol, uint8_t *__restrict__ sel, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sel[i] = support[i] && hypot(ux[i] - ax[i], uy[i] - ay[i]) <= 3.0 * tol;
}

__global__ void k_gate_final(const double *__restrict__ s1, const uint8_t *__restrict__ ok_fit,
                             const double *__restrict__ ax, const double *__restrict__ ay,
                             const double *__restrict__ tx, const double *__restrict__ ty, double min_support,
                             double tol, uint8_t *__restrict__ ok, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ok[i] = s1[i] + 0.5 >= min_support && ok_fit[i] && hypot(tx[i] - ax[i], ty[i] - ay[i]) <= tol;
}

// the gate's inputs from the frame state: support = good & roi, u = the direct value there
__global__ void k_gate_support(const uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask,
                               const double *__restrict__ cx, const double *__restrict__ cy,
                               uint8_t *__restrict__ support, double *__restrict__ gx, double *__restrict__ gy,
                               size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool good = (flags[i] & HYB_GOOD) != 0;
    support[i] = good && in_roi(mask, i);
    gx[i] = good ? cx[i] : NAN;
    gy[i] = good ? cy[i] : NAN;
}

__global__ void k_good_of(const uint8_t *__restrict__ flags, uint8_t *__restrict__ good, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) good[i] = (flags[i] & HYB_GOOD) != 0;
}

__global__ void k_zero_ints(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

void hyb_zero_counters(int *counters, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, counters, HYB_NCOUNTERS);
    LAUNCHCHECK();
}

void hyb_good_of(const uint8_t *flags, uint8_t *good, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_good_of, dim3(nblk(n)), dim3(256), 0, s, flags, good, n);
    LAUNCHCHECK();
}

void hyb_refmove(const uint8_t *good_prev, const double *cp, double *c0r, uint8_t *anch, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_refmove, dim3(nblk(n)), dim3(256), 0, s, good_prev, cp, c0r, anch, n);
    LAUNCHCHECK();
}

void hyb_classify(const double *cx, const double *cy, const double *leg_c, const double *c0r, const uint8_t *anch,
                  const uint8_t *mask, const double *upx, const double *upy, const int *nfb, double conf_min,
                  int fb_max, int reacq_ok, double *link, uint8_t *flags, int *counters, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_classify, dim3(nblk(n)), dim3(256), 0, s, cx, cy, leg_c, c0r, anch, mask, upx, upy, nfb,
                       conf_min, fb_max, reacq_ok, link, flags, counters, n);
    LAUNCHCHECK();
}

void hyb_carried(const double *cax, const double *cay, const double *upx, const double *upy, double *ax, double *ay,
                 size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_carried, dim3(nblk(n)), dim3(256), 0, s, cax, cay, upx, upy, ax, ay, n);
    LAUNCHCHECK();
}

void hyb_has_fb(uint8_t *flags, const double *ca, double conf_min, const uint8_t *mask, int *counters, size_t n,
                hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_has_fb, dim3(nblk(n)), dim3(256), 0, s, flags, ca, conf_min, mask, counters, n);
    LAUNCHCHECK();
}

void hyb_reacquire(uint8_t *flags, const uint8_t *gate0, const double *cx, const double *cy, const double *ax,
                   const double *ay, const double *leg_c, const double *ca, double fb_tol, double conf_min, int fb_max,
                   double *c0r, uint8_t *anch, int *nfb, const uint8_t *mask, int *counters, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_reacquire, dim3(nblk(n)), dim3(256), 0, s, flags, gate0, cx, cy, ax, ay, leg_c, ca, fb_tol,
                       conf_min, fb_max, c0r, anch, nfb, mask, counters, n);
    LAUNCHCHECK();
}

void hyb_output(const uint8_t *flags, const uint8_t *gate, const double *cx, const double *cy, const double *ax,
                const double *ay, const double *link, const double *ca, const int *nfb, int fb_max,
                const uint8_t *mask, double *ux, double *uy, float *conf, uint8_t *low, int32_t *fb, int *counters,
                size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_hyb_output, dim3(nblk(n)), dim3(256), 0, s, flags, gate, cx, cy, ax, ay, link, ca, nfb, fb_max,
                       mask, ux, uy, conf, low, fb, counters, n);
    LAUNCHCHECK();
}

size_t gate_scratch_bytes(int w, int h) {
    const size_t n = (size_t)w * h;
    return n * (12 + 5) * sizeof(double) + n * 4;
}

void hyb_gate(const uint8_t *flags, const uint8_t *mask, const double *cx, const double *cy, const double *tx,
              const double *ty, int w, int h, double radius, long min_support, double tol, void *scratch,
              uint8_t *ok, hipStream_t s) {
    const size_t n = (size_t)w * h;
    double *rows = (double *)scratch;
    double *gx = rows + 12 * n, *gy = gx + n, *ax = gy + n, *ay = ax + n, *s1 = ay + n;
    uint8_t *support = (uint8_t *)(s1 + n), *sel = support + n, *ok_fit = sel + n;
    const int r = (int)std::max(1.0, std::nearbyint(radius));
    const double ridge = 1e-3 * (double)r * (double)r;
    hipLaunchKernelGGL(k_gate_support, dim3(nblk(n)), dim3(256), 0, s, flags, mask, cx, cy, support, gx, gy, n);
    dcopy(sel, support, n, s);
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_gate_rows, grid2(w, h), dim3(256), 0, s, sel, support, gx, gy, w, h, r, rows);
        hipLaunchKernelGGL(k_gate_cols_fit, dim3((w + 255) / 256, (h + GATE_SEG - 1) / GATE_SEG), dim3(256), 0, s,
                           rows, w, h, r, ridge, ax, ay, s1, ok_fit);
        if (pass == 0)
            hipLaunchKernelGGL(k_gate_reselect, dim3(nblk(n)), dim3(256), 0, s, support, gx, gy, ax, ay, tol, sel, n);
        LAUNCHCHECK();
    }
    hipLaunchKernelGGL(k_gate_final, dim3(nblk(n)), dim3(256), 0, s, s1, ok_fit, ax, ay, tx, ty, (double)min_support, tol,
                       ok, n);
    LAUNCHCHECK();
}

}  // namespace ofx
