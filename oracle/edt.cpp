// Exact Euclidean distance transform, feature transform and 4-connected component filter for
// the oracle (test infrastructure only). Semantics of scipy.ndimage.distance_transform_edt and
// scipy.ndimage.label as the reference calls them (python/optflow_ref/phaseflow.py:346-348,
// track.py:45-53, track.py:86); the tie rule of the feature transform (smallest column, then
// smallest row) is the one rust/crates/optflow-core/src/edt.rs:175-183 documents for scipy
// 1.18.1: the pass along y runs first and both passes keep the smaller position on a tie.
#include <algorithm>
#include <limits>

#include "common.hpp"

namespace orc {

namespace {
const double INF = std::numeric_limits<double>::infinity();

// 1-D lower envelope of parabolas (q - p)^2 + f[p] over finite f; writes the minimum value and
// the minimising position (smaller p on a tie). Returns false when f has no finite entry.
bool envelope_1d(const double *f, int n, double *d2, int *arg, std::vector<int> &v,
                 std::vector<double> &z) {
    int k = -1;
    for (int q = 0; q < n; q++) {
        if (std::isinf(f[q])) continue;
        if (k < 0) {
            k = 0;
            v[0] = q;
            z[0] = -INF;
            z[1] = INF;
            continue;
        }
        const double fq = f[q] + (double)q * q;
        while (true) {
            const int p = v[k];
            const double s = (fq - (f[p] + (double)p * p)) / (2.0 * (q - p));
            if (s <= z[k]) {
                k--;
            } else {
                k++;
                v[k] = q;
                z[k] = s;
                z[k + 1] = INF;
                break;
            }
        }
    }
    if (k < 0) return false;
    int j = 0;
    for (int q = 0; q < n; q++) {
        while (z[j + 1] < q) j++;
        const double d = q - v[j];
        d2[q] = d * d + f[v[j]];
        if (arg) arg[q] = v[j];
    }
    return true;
}
}  // namespace

void edt(const uint8_t *input, int h, int w, double *dist, int32_t *nearest) {
    // pass 1 along y (per column), stored transposed [x * h + y]
    std::vector<double> cd2((size_t)w * h);
    std::vector<int> crow((size_t)w * h, 0);
    bool any_zero = false;
    for (size_t i = 0; i < (size_t)w * h; i++)
        if (!input[i]) { any_zero = true; break; }
#pragma omp parallel
    {
        std::vector<double> f(h), d2(h);
        std::vector<int> arg(h), v(h + 1);
        std::vector<double> z(h + 2);
#pragma omp for schedule(static)
        for (int x = 0; x < w; x++) {
            for (int y = 0; y < h; y++) f[y] = input[(size_t)y * w + x] ? INF : 0.0;
            if (!envelope_1d(f.data(), h, d2.data(), arg.data(), v, z)) {
                for (int y = 0; y < h; y++) { cd2[(size_t)x * h + y] = INF; crow[(size_t)x * h + y] = 0; }
            } else {
                for (int y = 0; y < h; y++) { cd2[(size_t)x * h + y] = d2[y]; crow[(size_t)x * h + y] = arg[y]; }
            }
        }
    }
    // pass 2 along x (per row)
#pragma omp parallel
    {
        std::vector<double> f(w), d2(w);
        std::vector<int> arg(w), v(w + 1);
        std::vector<double> z(w + 2);
#pragma omp for schedule(static)
        for (int y = 0; y < h; y++) {
            for (int x = 0; x < w; x++) f[x] = cd2[(size_t)x * h + y];
            bool ok = envelope_1d(f.data(), w, d2.data(), arg.data(), v, z);
            for (int x = 0; x < w; x++) {
                size_t i = (size_t)y * w + x;
                if (!any_zero || !ok) {
                    // scipy's answer without any background pixel (edt.rs:141-147)
                    if (dist) dist[i] = std::sqrt((double)(y + 1) * (y + 1) + (double)x * x);
                    if (nearest) nearest[i] = -1;
                } else {
                    if (dist) dist[i] = std::sqrt(d2[x]);
                    if (nearest) nearest[i] = crow[(size_t)arg[x] * h + y] * w + arg[x];
                }
            }
        }
    }
}

void large_components(const uint8_t *b, int h, int w, double min_area, uint8_t *out) {
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> seen(n, 0);
    std::vector<size_t> comp;
    std::fill(out, out + n, 0);
    for (size_t s = 0; s < n; s++) {
        if (!b[s] || seen[s]) continue;
        seen[s] = 1;
        comp.clear();
        comp.push_back(s);
        for (size_t i = 0; i < comp.size(); i++) {
            size_t p = comp[i];
            int x = (int)(p % w), y = (int)(p / w);
            size_t nb[4];
            int cnt = 0;
            if (x > 0) nb[cnt++] = p - 1;
            if (x + 1 < w) nb[cnt++] = p + 1;
            if (y > 0) nb[cnt++] = p - w;
            if (y + 1 < h) nb[cnt++] = p + w;
            for (int j = 0; j < cnt; j++)
                if (b[nb[j]] && !seen[nb[j]]) { seen[nb[j]] = 1; comp.push_back(nb[j]); }
        }
        if ((double)comp.size() >= min_area)
            for (size_t p : comp) out[p] = 1;
    }
}

}  // namespace orc

extern "C" void orc_edt(const uint8_t *input, int h, int w, double *dist, int32_t *nearest) {
    orc::edt(input, h, w, dist, nearest);
}

extern "C" void orc_large_components(const uint8_t *b, int h, int w, double min_area, uint8_t *out) {
    orc::large_components(b, h, w, min_area, out);
}
