// Cubic B-spline sampling on the device: the clamp, mirrored taps and weights of
// python/optflow_ref/interp.py:37-69 (rust/crates/optflow-core/src/interp.rs:77-106). Positions
// are double (a float cannot hold x = 2000.123 to better than 1.2e-4 px); weights and the
// coefficient sums are float, the precision the coefficients are stored in.
#pragma once
#include <hip/hip_runtime.h>

namespace ofx {

__device__ __forceinline__ int mirror_idx(int i, int n) {
    if (n == 1) return 0;
    const int p = 2 * (n - 1);
    i = abs(i) % p;
    return i >= n ? p - i : i;
}

struct Taps {
    int idx[4];
    float w[4];
};

// taps along one axis of length n at coordinate c (clamped to [0, n - 1])
__device__ __forceinline__ Taps taps_at(double c, int n) {
    c = fmin(fmax(c, 0.0), (double)(n - 1));
    const double f = floor(c);
    const int i0 = (int)f;
    const float t = (float)(c - f);
    const float t2 = t * t, t3 = t2 * t, u = 1.f - t;
    Taps T;
    T.w[0] = u * u * u * (1.f / 6.f);
    T.w[1] = (3.f * t3 - 6.f * t2 + 4.f) * (1.f / 6.f);
    T.w[2] = (-3.f * t3 + 3.f * t2 + 3.f * t + 1.f) * (1.f / 6.f);
    T.w[3] = t3 * (1.f / 6.f);
    if (i0 >= 1 && i0 + 2 < n) {
#pragma unroll
        for (int k = 0; k < 4; k++) T.idx[k] = i0 + k - 1;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) T.idx[k] = mirror_idx(i0 + k - 1, n);
    }
    return T;
}

__device__ __forceinline__ float sample_real(const float *__restrict__ c, int w, const Taps &tx, const Taps &ty) {
    float out = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float *row = c + (size_t)ty.idx[j] * w;
        float acc = tx.w[0] * row[tx.idx[0]];
#pragma unroll
        for (int k = 1; k < 4; k++) acc = fmaf(tx.w[k], row[tx.idx[k]], acc);
        out = fmaf(ty.w[j], acc, out);
    }
    return out;
}

__device__ __forceinline__ float2 sample_cplx(const float2 *__restrict__ c, int w, const Taps &tx, const Taps &ty) {
    float ox = 0.f, oy = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float2 *row = c + (size_t)ty.idx[j] * w;
        float ax = 0.f, ay = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float2 v = row[tx.idx[k]];
            ax = fmaf(tx.w[k], v.x, ax);
            ay = fmaf(tx.w[k], v.y, ay);
        }
        ox = fmaf(ty.w[j], ax, ox);
        oy = fmaf(ty.w[j], ay, oy);
    }
    return make_float2(ox, oy);
}

}  // namespace ofx
