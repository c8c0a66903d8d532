// Internal helpers of the FP64 oracle (test infrastructure only; see oracle.h).
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "oracle.h"

namespace orc {

using cd = std::complex<double>;
using Vec = std::vector<double>;

constexpr int PAD = 32;                 // python/optflow_ref/phaseflow.py:47
constexpr double APERTURE_TOL = 1e-3;   // phaseflow.py:50
constexpr double RIDGE = 1e-12;         // phaseflow.py:153

void set_error(const std::string &msg);

// FFT of one complex line in place (length n, any n).
void fft_line(cd *x, int n, bool inverse);
// 2-D FFT in place of an h x w complex array (unnormalised forward, 1/(h w) inverse).
void fft2(std::vector<cd> &a, int h, int w, bool inverse);

// interp
std::vector<double> gauss_kernel(double sigma);
void bspline_real(std::vector<double> &c, int h, int w);
void bspline_complex(std::vector<cd> &c, int h, int w);
int mirror(long i, int n);
void weights(double t, double wv[4]);
template <class T>
T sample_at(const T *coeffs, int h, int w, double x, double y);
enum Mode { NEAREST, CONSTANT };
// scipy.ndimage.correlate1d along axis (0 = rows direction y, 1 = x)
void correlate1d(const double *in, int h, int w, const double *k, int klen, int axis, Mode mode,
                 double *out);
void gauss_blur(const double *in, int h, int w, double sigma, double *out);

// numpy median (NaN if any NaN or empty)
double median(std::vector<double> v);

// EDT: squared-distance-free Euclidean distance to the nearest zero of `input` (nonzero =
// foreground), optional nearest-zero index (scipy rule).
void edt(const uint8_t *input, int h, int w, double *dist, int32_t *nearest);
void large_components(const uint8_t *b, int h, int w, double min_area, uint8_t *out);

// coarse stage (poc.cpp): grid positions along one axis; the dense start of track.py:131-137
// (gx_out, gy_out: h x w, or NULL) and the level-0 grid (ny x nx, or NULL)
std::vector<int> grid_positions(int n, int win, int step);
void coarse_dense(const double *I1, const double *I2, int H, int W, const orc_cfg &c, const double *init_x,
                  const double *init_y, double *gx_out, double *gy_out, std::vector<double> *grid_ux,
                  std::vector<double> *grid_uy, std::vector<double> *grid_peak);

// fine stage (flow.cpp): refine of an image-2 window (phaseflow.rs:1441-1456) and the confidence
// of a whole field
void refine_shifted(const double *i1, int h, int w, const double *i2, int h2, int w2, long shx, long shy,
                    const double *ux0, const double *uy0, const orc_cfg &c, double *ux_out, double *uy_out,
                    double *conf_out, double *resid_out, double *energy_out);
void confidence_of(const double *energy, const double *resid, const double *ux, const double *uy, int h, int w,
                   const orc_cfg &c, double *conf_out);

struct Cfg {
    const orc_cfg &c;
    explicit Cfg(const orc_cfg &cc) : c(cc) {}
    double max_scale() const {
        double m = c.scales[0];
        for (int i = 1; i < c.n_scales; i++) m = std::max(m, c.scales[i]);
        return m;
    }
    double auto_edge_guard_px() const {
        return 3.0 * c.lam * max_scale() / (2 * M_PI * c.bandwidth);
    }
    double edge_guard_px() const { return c.edge_guard >= 0 ? c.edge_guard : auto_edge_guard_px(); }
    double edge_guard_min_area_px() const {
        if (c.edge_guard_min_area >= 0) return c.edge_guard_min_area;
        double l = c.lam * max_scale();
        return 2.0 * l * l;
    }
    double fb_radius_px() const { return c.fb_radius > 0 ? c.fb_radius : 4.0 * c.sigma_w; }
    int n_filters() const { return c.n_orient * c.n_scales; }
};

}  // namespace orc
