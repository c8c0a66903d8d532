// The output side of the path on the host: the automatic ROI crop of the reference CLI and the
// writer of a result folder in the reference's format (.npy maps, timing.json, crop.json). No
// device work; the Python package and the command-line twin both write through these functions,
// so a folder is the same bytes whichever of them wrote it.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace ofx {
namespace {

constexpr int PAD_PX = 32;  // the response padding of the fine stage (phaseflow.py:47)

// ------------------------------------------------------------------ crop geometry
// The margin the reference CLI puts around the ROI by default (crop.rs:36-63): the largest
// displacement, the fine stage's window (4 sigma_w), the guard band, the response padding and the
// coarse stage's footprint (poc_win / 2 + grid_step) at full resolution, twice that for a
// sequence. Saturates instead of overflowing.
int64_t auto_margin(const ofx_config &c, int w, int h, bool seq) {
    const int L = pyramid_levels(c.max_disp, c.poc_win, h, w);
    const double parts[5] = {std::ceil(c.max_disp), std::ceil(4.0 * c.sigma_w), std::ceil(guard_band_px(c)),
                             (double)PAD_PX, (double)(c.poc_win / 2 + c.grid_step) * std::ldexp(1.0, L)};
    double total = 0;
    for (double v : parts) total += std::max(v, 0.0);
    if (seq) total *= 2;
    const double cap = (double)std::numeric_limits<int64_t>::max() / 4;
    return (int64_t)std::min(total, cap);
}

// ------------------------------------------------------------------ number spelling
// A double as serde_json writes it (the reference CLI's timing.json): the shortest digits that
// read back to the same value, laid out as the ryu crate does -- plain decimals with ".0" for
// integers while the decimal point sits within 16 digits and the value is not below 1e-5 or so,
// otherwise d.ddde[-]x. Non-finite values are written as null.
std::string json_number(double v) {
    if (!std::isfinite(v)) return "null";
    if (v == 0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; prec++) {
        std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    // buf: [-]d.ddde(+|-)xx with prec significant digits
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t epos = s.find('e');
    std::string digits;
    for (size_t i = 0; i < epos; i++)
        if (s[i] != '.') digits += s[i];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int exp10 = std::atoi(s.c_str() + epos + 1);
    const int len = (int)digits.size();
    const int kk = exp10 + 1;  // position of the decimal point after the first kk digits
    std::string out;
    if (kk >= len && kk <= 16) {
        out = digits + std::string(kk - len, '0') + ".0";
    } else if (kk > 0 && kk <= 16) {
        out = digits.substr(0, kk) + "." + digits.substr(kk);
    } else if (kk > -5 && kk <= 0) {
        out = "0." + std::string(-kk, '0') + digits;
    } else {
        out = digits.substr(0, 1);
        if (len > 1) out += "." + digits.substr(1);
        out += "e" + std::to_string(kk - 1);
    }
    return neg ? "-" + out : out;
}

// ------------------------------------------------------------------ files
void make_dirs(const std::string &dir) {
    if (dir.empty()) return;
    std::string part;
    size_t pos = 0;
    while (pos != std::string::npos) {
        pos = dir.find('/', pos + 1);
        part = dir.substr(0, pos);
        if (part.empty()) continue;
        if (::mkdir(part.c_str(), 0777) != 0 && errno != EEXIST)
            throw Error(OFX_EINVAL, "cannot create " + part + ": " + std::strerror(errno));
    }
    struct stat st;
    if (::stat(dir.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) throw Error(OFX_EINVAL, "cannot create " + dir);
}

struct File {
    FILE *f = nullptr;
    std::string path;
    explicit File(const std::string &p) : path(p) {
        f = std::fopen(p.c_str(), "wb");
        if (!f) throw Error(OFX_EINVAL, "cannot write " + p + ": " + std::strerror(errno));
    }
    void put(const void *data, size_t bytes) {
        if (bytes && std::fwrite(data, 1, bytes, f) != bytes) throw Error(OFX_EINVAL, "cannot write " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

// The header numpy.save writes for a C-ordered 2-D array (format 1.0): the dict text, spare room
// for the first axis to grow to 21 digits, spaces and a newline up to a multiple of 64 bytes.
std::string npy_header(const char *descr, int h, int w) {
    const std::string first = std::to_string(h);
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" + first + ", " +
                       std::to_string(w) + "), }";
    dict += std::string(21 - first.size(), ' ');
    const size_t body = dict.size() + 1;          // with the closing newline
    const size_t pad = 64 - (10 + body) % 64;     // magic 6 + version 2 + length 2
    const size_t hlen = body + pad;
    std::string out("\x93NUMPY\x01\x00", 8);
    out += (char)(hlen & 0xFF);
    out += (char)(hlen >> 8);
    out += dict + std::string(pad, ' ') + "\n";
    return out;
}

// Rows of a map written with a crop are formed in a buffer of about this many bytes and written in
// one call each; a map without a crop is written in one call.
constexpr size_t ROW_CHUNK = 4u << 20;

// One map of the result (w x h) as the map written (fw x fh): the crop's rectangle holds it, the
// rest is `fill`; the whole map when there is no crop.
template <class T>
void write_map(const std::string &path, const char *descr, const T *data, int w, int h, const ofx_crop *crop,
               int fw, int fh, T fill) {
    File f(path);
    const int ow = crop ? fw : w, oh = crop ? fh : h;
    const std::string hdr = npy_header(descr, oh, ow);
    f.put(hdr.data(), hdr.size());
    if (!crop) {
        f.put(data, (size_t)w * h * sizeof(T));
        return;
    }
    const int rows = (int)std::max<size_t>(1, ROW_CHUNK / ((size_t)ow * sizeof(T)));
    std::vector<T> buf((size_t)std::min(rows, oh) * ow);
    for (int y0 = 0; y0 < oh; y0 += rows) {
        const int y1 = std::min(oh, y0 + rows);
        for (int y = y0; y < y1; y++) {
            T *row = buf.data() + (size_t)(y - y0) * ow;
            if (y >= crop->y0 && y < crop->y1) {
                std::fill(row, row + crop->x0, fill);
                std::memcpy(row + crop->x0, data + (size_t)(y - crop->y0) * w, (size_t)w * sizeof(T));
                std::fill(row + crop->x0 + w, row + ow, fill);
            } else {
                std::fill(row, row + ow, fill);
            }
        }
        f.put(buf.data(), (size_t)(y1 - y0) * ow * sizeof(T));
    }
}

// valid: 1 where the written ux is finite
void write_valid(const std::string &path, const float *ux, int w, int h, const ofx_crop *crop, int fw, int fh) {
    File f(path);
    const int ow = crop ? fw : w, oh = crop ? fh : h;
    const std::string hdr = npy_header("|u1", oh, ow);
    f.put(hdr.data(), hdr.size());
    const int rows = (int)std::max<size_t>(1, ROW_CHUNK / (size_t)ow);
    std::vector<uint8_t> buf((size_t)std::min(rows, oh) * ow);
    const int x0 = crop ? crop->x0 : 0, ys = crop ? crop->y0 : 0;
    for (int y0 = 0; y0 < oh; y0 += rows) {
        const int y1 = std::min(oh, y0 + rows);
        std::fill(buf.begin(), buf.end(), 0);
        for (int y = std::max(y0, ys); y < std::min(y1, ys + h); y++) {
            uint8_t *row = buf.data() + (size_t)(y - y0) * ow + x0;
            const float *src = ux + (size_t)(y - ys) * w;
            for (int x = 0; x < w; x++) row[x] = std::isfinite(src[x]) ? 1 : 0;
        }
        f.put(buf.data(), (size_t)(y1 - y0) * ow);
    }
}

// Runs the jobs on up to `threads` host threads (the calling one included): the files of a result
// folder are independent, and the page cache takes them faster side by side. The first error is
// rethrown once every job has ended.
void run_jobs(const std::vector<std::function<void()>> &jobs, unsigned threads) {
    std::atomic<size_t> next{0};
    std::mutex m;
    std::exception_ptr err;
    auto work = [&] {
        for (size_t j; (j = next.fetch_add(1)) < jobs.size();) {
            try {
                jobs[j]();
            } catch (...) {
                std::lock_guard<std::mutex> l(m);
                if (!err) err = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    const unsigned n = std::min<unsigned>(threads, (unsigned)jobs.size());
    for (unsigned t = 1; t < n; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (err) std::rethrow_exception(err);
}

// Host threads for writing one folder: the files (up to 12), capped by the cores present.
unsigned writer_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(6u, hc ? hc : 1u));
}

void write_text(const std::string &path, const std::string &text) {
    File f(path);
    f.put(text.data(), text.size());
}

}  // namespace
}  // namespace ofx

using namespace ofx;

extern "C" {

int ofx_pyramid_levels(const ofx_config *cfg, int w, int h, int *levels) {
    return guarded_call([&] {
        validate_config(cfg);
        if (!levels) throw Error(OFX_EINVAL, "levels is NULL");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty image");
        *levels = pyramid_levels(cfg->max_disp, cfg->poc_win, h, w);
    });
}

int ofx_poc_level_window(int w, int h, int win, int *wy, int *wx) {
    return guarded_call([&] {
        if (!wy || !wx) throw Error(OFX_EINVAL, "wy or wx is NULL");
        if (w < 1 || h < 1 || win < 1) throw Error(OFX_EINVAL, "empty image or window");
        int a = 0, b = 0;
        if (poc_window_at(h, w, win, a, b)) {
            *wy = a;
            *wx = b;
        } else {
            *wy = *wx = 0;
        }
    });
}

int ofx_crop_plan(const uint8_t *mask, int w, int h, const ofx_config *cfg, const int64_t *margin, int seq,
                  ofx_crop *out) {
    return guarded_call([&] {
        validate_config(cfg);
        if (!mask || !out) throw Error(OFX_EINVAL, "mask or out is NULL");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty mask");
        // bounding box of the ROI: rows scanned once, each row's first and last set column
        int bx0 = w, bx1 = -1, by0 = h, by1 = -1;
        for (int y = 0; y < h; y++) {
            const uint8_t *r = mask + (size_t)y * w;
            int a = 0;
            while (a < w && !r[a]) a++;
            if (a == w) continue;
            int b = w - 1;
            while (!r[b]) b--;
            bx0 = std::min(bx0, a);
            bx1 = std::max(bx1, b);
            if (by1 < 0) by0 = y;
            by1 = y;
        }
        if (bx1 < 0) throw Error(OFX_EINVAL, "ROI mask is empty");
        if (margin && *margin < 0) throw Error(OFX_EINVAL, "--crop-margin must be >= 0, got " + std::to_string(*margin));
        if (std::min(w, h) < cfg->poc_win)
            throw Error(OFX_EINVAL, "image smaller than poc_win (" + std::to_string(cfg->poc_win) + "): (" +
                                        std::to_string(h) + ", " + std::to_string(w) + ")");
        const int64_t m = margin ? *margin : auto_margin(*cfg, w, h, seq != 0);
        // the crop's origin on the coarse grid of the frame: a multiple of grid_step 2^L
        const int64_t step = (int64_t)cfg->grid_step << pyramid_levels(cfg->max_disp, cfg->poc_win, h, w);
        auto down = [&](int64_t v) { return std::max<int64_t>(v, 0) / step * step; };
        ofx_crop c;
        c.x0 = (int)down(bx0 - m);
        c.y0 = (int)down(by0 - m);
        c.x1 = (int)std::min<int64_t>((int64_t)bx1 + 1 + m, w);
        c.y1 = (int)std::min<int64_t>((int64_t)by1 + 1 + m, h);
        c.margin = m;
        const int cw = c.x1 - c.x0, ch = c.y1 - c.y0;
        if (std::min(cw, ch) < cfg->poc_win)
            throw Error(OFX_EINVAL, "the crop " + std::to_string(cw) + "x" + std::to_string(ch) +
                                        " is smaller than poc_win (" + std::to_string(cfg->poc_win) +
                                        "): increase --crop-margin or use --crop off");
        *out = c;
    });
}

int ofx_write_crop_json(const char *dir, const ofx_crop *crop) {
    return guarded_call([&] {
        if (!dir) throw Error(OFX_EINVAL, "dir is NULL");
        make_dirs(dir);
        std::string text;
        if (!crop) {
            text = "{\n  \"mode\": \"off\"\n}";
        } else {
            text = "{\n  \"x0\": " + std::to_string(crop->x0) + ",\n  \"y0\": " + std::to_string(crop->y0) +
                   ",\n  \"x1\": " + std::to_string(crop->x1) + ",\n  \"y1\": " + std::to_string(crop->y1) +
                   ",\n  \"margin\": " + std::to_string(crop->margin) + ",\n  \"mode\": \"auto\"\n}";
        }
        write_text(std::string(dir) + "/crop.json", text);
    });
}

int ofx_write_result(const char *dir, const ofx_result *res, int w, int h, const ofx_crop *crop, int frame_w,
                     int frame_h, int raw, int seq) {
    return guarded_call([&] {
        if (!dir || !res) throw Error(OFX_EINVAL, "dir or result is NULL");
        if (w < 1 || h < 1) throw Error(OFX_EINVAL, "empty result");
        if (crop) {
            if (crop->x0 < 0 || crop->y0 < 0 || crop->x1 > frame_w || crop->y1 > frame_h ||
                crop->x1 - crop->x0 != w || crop->y1 - crop->y0 != h)
                throw Error(OFX_EINVAL, "the crop " + std::to_string(crop->x1 - crop->x0) + "x" +
                                            std::to_string(crop->y1 - crop->y0) + " at (" +
                                            std::to_string(crop->x0) + ", " + std::to_string(crop->y0) +
                                            ") does not hold a " + std::to_string(w) + "x" + std::to_string(h) +
                                            " result inside a " + std::to_string(frame_w) + "x" +
                                            std::to_string(frame_h) + " frame");
        }
        const struct {
            const char *name;
            const float *data;
            float fill;
        } maps[8] = {{"ux", res->ux, NAN},   {"uy", res->uy, NAN},   {"conf", res->conf, 0.0f},
                     {"Exx", res->exx, NAN}, {"Eyy", res->eyy, NAN}, {"Exy", res->exy, NAN},
                     {"gmax", res->gmax, NAN}, {"vol", res->vol, NAN}};
        for (const auto &m : maps)
            if (!m.data) throw Error(OFX_EINVAL, std::string("result map ") + m.name + " is NULL");
        if (raw && (!res->raw_ux || !res->raw_uy)) throw Error(OFX_EINVAL, "raw maps requested but NULL");
        make_dirs(dir);
        const std::string d(dir);
        std::vector<std::function<void()>> jobs;
        for (const auto &m : maps)
            jobs.push_back([&, m] {
                write_map<float>(d + "/" + m.name + ".npy", "<f4", m.data, w, h, crop, frame_w, frame_h, m.fill);
            });
        jobs.push_back([&] { write_valid(d + "/valid.npy", res->ux, w, h, crop, frame_w, frame_h); });
        if (res->fb)
            jobs.push_back([&] { write_map<int32_t>(d + "/fb.npy", "<i4", res->fb, w, h, crop, frame_w, frame_h, 0); });
        if (raw) {
            jobs.push_back(
                [&] { write_map<float>(d + "/raw_ux.npy", "<f4", res->raw_ux, w, h, crop, frame_w, frame_h, NAN); });
            jobs.push_back(
                [&] { write_map<float>(d + "/raw_uy.npy", "<f4", res->raw_uy, w, h, crop, frame_w, frame_h, NAN); });
        }
        run_jobs(jobs, writer_threads());
        // timing.json in the reference CLI's layout: its Timing fields in their order, "device"
        // after the first, the sequence fields only for sequences, the hybrid shares only with fb
        std::string t = "{\n  \"coarse\": " + json_number(res->t_coarse) + ",\n  \"device\": \"gpu\",\n  \"fine\": " +
                        json_number(res->t_fine) + ",\n  \"strain\": " + json_number(res->t_strain);
        if (seq) t += ",\n  \"reference\": " + std::to_string(res->reference) + ",\n  \"mean_conf\": " +
                      json_number(res->mean_conf);
        if (seq && res->fb)
            t += ",\n  \"n_fb_max\": " + std::to_string(res->n_fb_max) + ",\n  \"fallback_frac\": " +
                 json_number(res->fallback_frac) + ",\n  \"fb_gate_rejected\": " + json_number(res->fb_gate_rejected) +
                 ",\n  \"reacquired\": " + json_number(res->reacquired);
        t += "\n}";
        write_text(d + "/timing.json", t);
    });
}

int ofx_format_json_number(double v, char *buf, size_t cap) {
    return guarded_call([&] {
        if (!buf || cap == 0) throw Error(OFX_EINVAL, "buf is NULL or empty");
        const std::string s = json_number(v);
        if (s.size() + 1 > cap) throw Error(OFX_EINVAL, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

}  // extern "C"
