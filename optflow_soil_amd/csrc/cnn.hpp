// The learned coarse proposal (cnn.hip): shapes, device state and host entry points.
#pragma once
#include "internal.hpp"

namespace ofx {

// correlation volume: displacements of -CV_R .. CV_R cells on both axes
constexpr int CV_R = 6;
constexpr int CV_D = 2 * CV_R + 1;
constexpr int CV_CH = CV_D * CV_D;  // 169
constexpr int CV_CPAD = 176;        // stored channels (a multiple of 16; the rest are 0)
constexpr int NET_CH = 32;          // hidden channels
constexpr int OUT_CPAD = 16;        // stored channels of the output gradient (2 used)
constexpr int NET_NLAYERS = 6;
constexpr int TRAIN_BUMPS = 6;

struct LayerShape {
    int cin, cin_pad, cout;
    bool relu;
};
extern const LayerShape NET_LAYERS[NET_NLAYERS];

// number of parameters in the flat layout (per layer: weights [cout][9][cin], then bias[cout])
size_t net_param_count();
// floats of the padded master layout (per layer: weights [32][9][cin_pad], then bias[32])
size_t net_master_floats();
size_t layer_master_floats(int l);
void net_flat_to_master(const float *flat, std::vector<float> &master);
void net_master_to_flat(const std::vector<float> &master, float *flat);

// a device buffer that grows to the largest size asked for and is freed with its owner
struct GrowBuf {
    void *p = nullptr;
    size_t bytes = 0;
    GrowBuf() = default;
    GrowBuf(const GrowBuf &) = delete;
    GrowBuf &operator=(const GrowBuf &) = delete;
    ~GrowBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T *get(size_t n) {
        const size_t b = std::max<size_t>(n, 1) * sizeof(T);
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

// The network's parameters on the device: float32 master copy (padded layout) and the bf16
// weights the forward and backward-by-data kernels read.
struct CoarseNet {
    float *master = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    size_t m_off[NET_NLAYERS] = {0}, wf_off[NET_NLAYERS] = {0}, wb_off[NET_NLAYERS] = {0};
    CoarseNet() = default;
    CoarseNet(const CoarseNet &) = delete;
    CoarseNet &operator=(const CoarseNet &) = delete;
    ~CoarseNet();
    void upload(const std::vector<float> &master_host, hipStream_t s);  // and cast
    void cast(hipStream_t s);                                           // master -> bf16 copies
    const float *bias(int l) const;
    static int bw_pad(int l);
};

// The activations of one pass over nb maps of h8 x w8 cells.
struct NetActs {
    int nb = 0, h8 = 0, w8 = 0;
    float *a1 = nullptr, *a2 = nullptr, *pred = nullptr;  // the two frames in cells; the output (cells)
    float2 *st = nullptr;                                  // mean, 1 / sd of each map (a1 maps, then a2)
    double *stp = nullptr;                                 // partial sums of the statistics
    uint16_t *cv = nullptr, *hid[NET_NLAYERS - 1] = {nullptr};
    GrowBuf a1b, a2b, stb, stpb, cvb, hb[NET_NLAYERS - 1], predb;
    void ensure(int nb, int h8, int w8);
};

// the correlation volume of A.a1, A.a2; then the six layers into A.pred
void net_features(NetActs &A, hipStream_t s);
void net_forward(const CoarseNet &net, NetActs &A, hipStream_t s);
// the proposal of a pair of frames (device, w x h): A.pred holds ((h + 7) / 8) x ((w + 7) / 8) x 2
// displacements in cells (1 cell = 8 px)
void net_propose(const CoarseNet &net, NetActs &A, const float *i1, const float *i2, int w, int h, hipStream_t s);
// the proposal at the POC grid x0 + j step, y0 + i step (nx x ny), in pixels
void prop_to_grid(const NetActs &A, int x0, int y0, int step, int nx, int ny, double *ix, double *iy, hipStream_t s);

// One training sample: where its crop lies in which pattern, its field (TRAIN_BUMPS Gaussian
// bumps in the crop's coordinates: u = sum a_k exp(-|X - c_k|^2 / (2 s_k^2))) and its noise seed.
struct TrainSample {
    int pattern, ox, oy;
    uint32_t seed;
    float cx[TRAIN_BUMPS], cy[TRAIN_BUMPS], inv2s2[TRAIN_BUMPS], ax[TRAIN_BUMPS], ay[TRAIN_BUMPS];
};
void make_train_samples(uint64_t seed, int first, int count, int n_pat, int pw, int ph, int crop, double max_disp,
                        std::vector<TrainSample> &out);
void render_train_pairs(const float *pats, int pw, int ph, const TrainSample *smp, int nb, int h8, int w8, float noise,
                        float *a1, float *a2, float *target, hipStream_t s);

struct NetTrainer {
    NetActs A;
    float *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *target = nullptr, *part = nullptr, *bpart = nullptr;
    uint16_t *dout = nullptr, *dz[2] = {nullptr, nullptr};
    double *lpart = nullptr;
    int ng = 0, lblk = 0;
    GrowBuf gradb, m1b, m2b, targetb, doutb, dzb[2], partb, bpartb, lpartb;
    PinnedBuf pin;
    void ensure(int nb, int h8, int w8);
    void zero_moments(hipStream_t s);
    void forward_backward(const CoarseNet &net, double *loss_dev, hipStream_t s);
    void adam_step(CoarseNet &net, int step, double lr, hipStream_t s);
};

}  // namespace ofx
