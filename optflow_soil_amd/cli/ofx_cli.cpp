// ofx_cli: the command-line twin of the reference's `optflow pair` / `optflow seq`
// (rust/crates/optflow-cli/src/args.rs:7-31, run.rs:164-207, :583-700) over the C ABI of libofx.so.
// It reads frames, applies the settings (--config JSON, --set K=V), plans the automatic ROI crop,
// tracks on the GPU and writes the reference's output folder: config.json, crop.json and the result
// maps (a pair into --out, sequence frame k into --out/frame_0000k), through the library's writer.
//
//   ofx_cli pair --ref FILE --def FILE [common options]
//   ofx_cli seq  --frames DIR|GLOB [--start N] [--stop N] [--step N] [common options]
//   common: --mask FILE  --scale N  --crop auto|off  --crop-margin N  --config FILE  --set K=V
//           --out DIR (required)  --raw  --device gpu
//
// Frames and masks are .npy (2-D uint8, uint16 or float32; uint8 / uint16 with 3 or 4 channels)
// or binary PGM (P5, 8 or 16 bit). Integer frames become grey levels v / 255 or v / 65535 (colour:
// 0.299 R + 0.587 G + 0.114 B) box-downscaled by --scale on the GPU (ofx_frame_to_gray); float32
// frames are taken as grey levels. Errors print "error: <message>" and exit with status 2, as the
// reference's CLI does.
#include <glob.h>
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <dirent.h>
#include <exception>
#include <limits.h>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/ofx.h"

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc) {
    if (rc != OFX_OK) throw Fail(ofx_last_error());
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// A bounded first-in first-out hand-off from one producing thread to one consuming thread. The
// producer ends the stream with finish() (passing its error, if any, to the consumer); the consumer
// gives up with abandon(), after which put() returns false at once.
template <class T>
class Channel {
  public:
    explicit Channel(size_t capacity) : cap_(capacity) {}
    bool put(T v) {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return q_.size() < cap_ || abandoned_; });
        if (abandoned_) return false;
        q_.push_back(std::move(v));
        cv_.notify_all();
        return true;
    }
    // false once the stream has ended and every item has been taken; rethrows the producer's error
    bool take(T &v) {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return !q_.empty() || finished_; });
        if (q_.empty()) {
            if (err_) std::rethrow_exception(err_);
            return false;
        }
        v = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return true;
    }
    void finish(std::exception_ptr err = nullptr) {
        std::lock_guard<std::mutex> l(m_);
        finished_ = true;
        err_ = err;
        cv_.notify_all();
    }
    void abandon() {
        std::lock_guard<std::mutex> l(m_);
        abandoned_ = true;
        q_.clear();
        cv_.notify_all();
    }

  private:
    size_t cap_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<T> q_;
    bool finished_ = false, abandoned_ = false;
    std::exception_ptr err_;
};

// ------------------------------------------------------------------ frames
struct Pixels {
    int w = 0, h = 0, channels = 1, bits = 8;  // bits 32: float grey levels
    std::vector<unsigned char> bytes;
};

std::string read_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Fail("cannot read " + path + ": " + std::strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

Pixels read_npy(const std::string &path) {
    const std::string s = read_file(path);
    if (s.size() < 10 || s.compare(0, 6, "\x93NUMPY") != 0) throw Fail(path + ": not a .npy file");
    const int major = (unsigned char)s[6];
    size_t hlen, off;
    if (major == 1) {
        hlen = (unsigned char)s[8] | ((size_t)(unsigned char)s[9] << 8);
        off = 10;
    } else {
        if (s.size() < 12) throw Fail(path + ": truncated header");
        hlen = 0;
        for (int i = 0; i < 4; i++) hlen |= (size_t)(unsigned char)s[8 + i] << (8 * i);
        off = 12;
    }
    if (s.size() < off + hlen) throw Fail(path + ": truncated header");
    const std::string hdr = s.substr(off, hlen);
    auto field = [&](const char *key) {
        const size_t k = hdr.find(key);
        if (k == std::string::npos) throw Fail(path + ": header lacks " + key);
        return hdr.substr(k + std::strlen(key));
    };
    std::string d = field("'descr':");
    d = d.substr(d.find('\'') + 1);
    d = d.substr(0, d.find('\''));
    if (field("'fortran_order':").find("True") < 3) throw Fail(path + ": Fortran-ordered arrays are not supported");
    std::string sh = field("'shape':");
    sh = sh.substr(sh.find('(') + 1);
    sh = sh.substr(0, sh.find(')'));
    std::vector<long> dims;
    for (char *p = &sh[0]; *p;) {
        char *e;
        const long v = std::strtol(p, &e, 10);
        if (e == p) {
            p++;
            continue;
        }
        dims.push_back(v);
        p = e;
    }
    Pixels px;
    if (dims.size() == 2) {
        px.h = (int)dims[0];
        px.w = (int)dims[1];
    } else if (dims.size() == 3 && (dims[2] == 3 || dims[2] == 4)) {
        px.h = (int)dims[0];
        px.w = (int)dims[1];
        px.channels = (int)dims[2];
    } else {
        throw Fail(path + ": expected an H x W (or H x W x 3|4) array");
    }
    size_t elem;
    if (d == "|u1" || d == "|b1") {
        px.bits = 8;
        elem = 1;
    } else if (d == "<u2") {
        px.bits = 16;
        elem = 2;
    } else if (d == "<f4") {
        px.bits = 32;
        elem = 4;
        if (px.channels != 1) throw Fail(path + ": float frames must be grey (H x W)");
    } else {
        throw Fail(path + ": unsupported dtype " + d + " (uint8, uint16 or float32)");
    }
    const size_t need = (size_t)px.w * px.h * px.channels * elem;
    if (s.size() < off + hlen + need) throw Fail(path + ": truncated data");
    px.bytes.assign(s.begin() + off + hlen, s.begin() + off + hlen + need);
    return px;
}

Pixels read_pgm(const std::string &path) {
    const std::string s = read_file(path);
    size_t p = 0;
    auto token = [&]() {
        while (p < s.size()) {
            if (s[p] == '#')
                while (p < s.size() && s[p] != '\n') p++;
            else if (std::isspace((unsigned char)s[p]))
                p++;
            else
                break;
        }
        const size_t a = p;
        while (p < s.size() && !std::isspace((unsigned char)s[p])) p++;
        return s.substr(a, p - a);
    };
    if (token() != "P5") throw Fail(path + ": only binary PGM (P5) is read");
    Pixels px;
    px.w = std::atoi(token().c_str());
    px.h = std::atoi(token().c_str());
    const int maxv = std::atoi(token().c_str());
    p++;  // the single whitespace before the raster
    if (px.w < 1 || px.h < 1 || maxv < 1 || maxv > 65535) throw Fail(path + ": bad PGM header");
    px.bits = maxv > 255 ? 16 : 8;
    const size_t n = (size_t)px.w * px.h * (px.bits / 8);
    if (s.size() < p + n) throw Fail(path + ": truncated data");
    px.bytes.assign(s.begin() + p, s.begin() + p + n);
    if (px.bits == 16)  // PGM is big-endian
        for (size_t i = 0; i + 1 < n; i += 2) std::swap(px.bytes[i], px.bytes[i + 1]);
    return px;
}

bool ends_with(const std::string &s, const char *suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

Pixels read_pixels(const std::string &path) {
    if (ends_with(path, ".npy")) return read_npy(path);
    if (ends_with(path, ".pgm")) return read_pgm(path);
    throw Fail(path + ": frames are read from .npy or .pgm files");
}

// grey levels of a frame (w / scale x h / scale)
std::vector<float> to_gray(ofx_engine *e, const Pixels &px, int scale, int &w, int &h) {
    if (px.bits == 32) {
        if (scale != 1) throw Fail("--scale needs integer frames (uint8 or uint16)");
        w = px.w;
        h = px.h;
        std::vector<float> out((size_t)w * h);
        std::memcpy(out.data(), px.bytes.data(), out.size() * sizeof(float));
        return out;
    }
    w = px.w / scale;
    h = px.h / scale;
    if (w < 1 || h < 1) throw Fail("--scale leaves an empty frame");
    std::vector<float> out((size_t)w * h);
    check(ofx_frame_to_gray(e, px.bytes.data(), px.bits, px.channels, px.w, px.h, scale, out.data()));
    return out;
}

// ROI mask: nonzero = inside; the shape of the (scaled) frame
std::vector<uint8_t> read_mask(const std::string &path, int w, int h) {
    const Pixels m = read_pixels(path);
    if (m.w != w || m.h != h || m.channels != 1)
        throw Fail("ROI mask shape (" + std::to_string(m.h) + ", " + std::to_string(m.w) +
                   ") differs from the image shape (" + std::to_string(h) + ", " + std::to_string(w) + ")");
    std::vector<uint8_t> out((size_t)w * h);
    const size_t step = m.bits / 8;
    for (size_t i = 0; i < out.size(); i++) {
        bool nz = false;
        for (size_t b = 0; b < step; b++) nz = nz || m.bytes[i * step + b] != 0;
        out[i] = nz;
    }
    return out;
}

std::vector<std::string> list_frames(const std::string &src) {
    std::vector<std::string> out;
    struct stat st;
    if (::stat(src.c_str(), &st) == 0 && S_ISDIR(st.st_mode)) {
        DIR *d = ::opendir(src.c_str());
        if (!d) throw Fail("cannot read " + src);
        while (dirent *e = ::readdir(d)) {
            const std::string n = e->d_name;
            if (ends_with(n, ".npy") || ends_with(n, ".pgm")) out.push_back(src + "/" + n);
        }
        ::closedir(d);
    } else {
        glob_t g;
        if (::glob(src.c_str(), 0, nullptr, &g) == 0)
            for (size_t i = 0; i < g.gl_pathc; i++) out.push_back(g.gl_pathv[i]);
        ::globfree(&g);
    }
    std::sort(out.begin(), out.end());
    if (out.empty()) throw Fail("no images found for " + src);
    return out;
}

// ------------------------------------------------------------------ settings
// A JSON value of a flat settings object: a number, a string, true / false, or a list of numbers.
struct Value {
    enum Kind { NUM, STR, BOOL, LIST } kind = NUM;
    double num = 0;
    bool integral = false, boolean = false;
    std::string str;
    std::vector<double> list;
};

struct Parser {
    const std::string &s;
    size_t p = 0;
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
    }
    [[noreturn]] void bad(const std::string &what) { throw Fail("config: " + what); }
    std::string string() {
        if (s[p] != '"') bad("expected a string");
        std::string out;
        for (p++; p < s.size() && s[p] != '"'; p++) {
            if (s[p] == '\\' && p + 1 < s.size()) p++;
            out += s[p];
        }
        if (p >= s.size()) bad("unterminated string");
        p++;
        return out;
    }
    double number(bool &integral) {
        const char *b = s.c_str() + p;
        char *e;
        const double v = std::strtod(b, &e);
        if (e == b) bad("expected a value");
        integral = std::string(b, (size_t)(e - b)).find_first_of(".eE") == std::string::npos;
        p += e - b;
        return v;
    }
    Value value() {
        ws();
        Value v;
        if (p >= s.size()) bad("expected a value");
        if (s[p] == '"') {
            v.kind = Value::STR;
            v.str = string();
        } else if (s.compare(p, 4, "true") == 0 || s.compare(p, 5, "false") == 0) {
            v.kind = Value::BOOL;
            v.boolean = s[p] == 't';
            p += v.boolean ? 4 : 5;
        } else if (s[p] == '[') {
            v.kind = Value::LIST;
            p++;
            ws();
            while (p < s.size() && s[p] != ']') {
                bool in;
                ws();
                v.list.push_back(number(in));
                ws();
                if (s[p] == ',') p++;
                ws();
            }
            if (p >= s.size()) bad("unterminated list");
            p++;
        } else {
            v.num = number(v.integral);
        }
        return v;
    }
    std::vector<std::pair<std::string, Value>> object() {
        std::vector<std::pair<std::string, Value>> out;
        ws();
        if (p >= s.size() || s[p] != '{') bad("config must be a JSON object");
        p++;
        ws();
        while (p < s.size() && s[p] != '}') {
            ws();
            const std::string k = string();
            ws();
            if (s[p] != ':') bad("expected ':'");
            p++;
            out.push_back({k, value()});
            ws();
            if (s[p] == ',') p++;
            ws();
        }
        return out;
    }
};

// The settings of a run: ofx_config plus the string keys of the reference's configuration.
struct Settings {
    ofx_config c;
    std::string seq_mode = "chain", coarse_model = "poc", tavg_model = "affine";
    int tavg = 0;
    bool tavg_align = false;
    Settings() { ofx_config_default(&c); }
};

// the keys in the order of the configuration (python/optflow_ref/config.py:24-90, then this
// engine's own: tile, coarse_model, coarse_gate)
const char *const KEYS[] = {"grid_step", "poc_win", "max_disp", "poc_peak_min", "median_thresh", "lam", "scales",
                            "n_orient", "bandwidth", "dc_cut", "sigma_w", "shape_order", "sigma_s", "outer_iters",
                            "inner_iters", "conf_min", "edge_guard", "edge_guard_min_area", "phase_tol", "strain_win",
                            "ref_update_conf", "seq_mode", "fb_max", "fb_radius", "fb_min_support", "fb_tol", "tavg",
                            "tavg_align", "tavg_model", "prewarp_freeze_at", "tile", "coarse_model", "coarse_gate"};

void set_key(Settings &st, const std::string &k, const Value &v) {
    ofx_config &c = st.c;
    auto num = [&]() {
        if (v.kind != Value::NUM) throw Fail("config key '" + k + "' must be a number");
        return v.num;
    };
    auto integer = [&]() {
        if (v.kind != Value::NUM || !v.integral) throw Fail("config key '" + k + "' must be an integer");
        return (int)v.num;
    };
    auto str = [&]() {
        if (v.kind != Value::STR) throw Fail("config key '" + k + "' must be a string");
        return v.str;
    };
    const std::map<std::string, int *> ints = {
        {"grid_step", &c.grid_step}, {"poc_win", &c.poc_win}, {"n_orient", &c.n_orient},
        {"shape_order", &c.shape_order}, {"outer_iters", &c.outer_iters}, {"inner_iters", &c.inner_iters},
        {"strain_win", &c.strain_win}, {"fb_max", &c.fb_max}, {"fb_min_support", &c.fb_min_support},
        {"prewarp_freeze_at", &c.prewarp_freeze_at}, {"tile", &c.tile}, {"tavg", &st.tavg}};
    const std::map<std::string, double *> dbls = {
        {"max_disp", &c.max_disp}, {"poc_peak_min", &c.poc_peak_min}, {"median_thresh", &c.median_thresh},
        {"lam", &c.lam}, {"bandwidth", &c.bandwidth}, {"dc_cut", &c.dc_cut}, {"sigma_w", &c.sigma_w},
        {"sigma_s", &c.sigma_s}, {"conf_min", &c.conf_min}, {"edge_guard", &c.edge_guard},
        {"edge_guard_min_area", &c.edge_guard_min_area}, {"phase_tol", &c.phase_tol},
        {"ref_update_conf", &c.ref_update_conf}, {"fb_radius", &c.fb_radius}, {"fb_tol", &c.fb_tol},
        {"coarse_gate", &c.coarse_gate}};
    if (ints.count(k)) {
        *ints.at(k) = integer();
    } else if (dbls.count(k)) {
        *dbls.at(k) = num();
    } else if (k == "scales") {
        if (v.kind != Value::LIST || v.list.empty() || v.list.size() > OFX_MAX_SCALES)
            throw Fail("config key 'scales' must be a list of numbers");
        c.n_scales = (int)v.list.size();
        for (size_t i = 0; i < v.list.size(); i++) c.scales[i] = v.list[i];
    } else if (k == "seq_mode") {
        st.seq_mode = str();
        if (st.seq_mode != "chain" && st.seq_mode != "hybrid")
            throw Fail("seq_mode must be 'chain' or 'hybrid', got '" + st.seq_mode + "'");
        c.seq_hybrid = st.seq_mode == "hybrid";
    } else if (k == "coarse_model") {
        st.coarse_model = str();
        if (st.coarse_model != "poc" && st.coarse_model != "learned")
            throw Fail("coarse_model must be 'poc' or 'learned', got '" + st.coarse_model + "'");
        c.coarse_learned = st.coarse_model == "learned";
    } else if (k == "tavg_model") {
        st.tavg_model = str();
    } else if (k == "tavg_align") {
        if (v.kind != Value::BOOL) throw Fail("config key 'tavg_align' must be true or false");
        st.tavg_align = v.boolean;
    } else {
        throw Fail("unknown config keys: ['" + k + "']");
    }
}

// v as Python's repr spells a float (json.dumps of the reference's config.json): the shortest
// digits that read back to v, plain decimals for exponents -4 .. 15, otherwise d.ddde+XX.
std::string py_float(double v) {
    if (v == 0) return std::signbit(v) ? "-0.0" : "0.0";
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    char buf[64];
    for (int prec = 1; prec <= 17; prec++) {
        std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t e = s.find('e');
    std::string digits;
    for (size_t i = 0; i < e; i++)
        if (s[i] != '.') digits += s[i];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int x = std::atoi(s.c_str() + e + 1);
    std::string out;
    if (x >= 16 || x < -4) {
        out = digits.substr(0, 1);
        if (digits.size() > 1) out += "." + digits.substr(1);
        char ex[16];
        std::snprintf(ex, sizeof ex, "e%c%02d", x < 0 ? '-' : '+', std::abs(x));
        out += ex;
    } else if (x < 0) {
        out = "0." + std::string(-x - 1, '0') + digits;
    } else if ((int)digits.size() <= x + 1) {
        out = digits + std::string(x + 1 - digits.size(), '0') + ".0";
    } else {
        out = digits.substr(0, x + 1) + "." + digits.substr(x + 1);
    }
    return neg ? "-" + out : out;
}

// config.json: every key in order, json.dumps(..., indent=2)
std::string config_json(const Settings &st) {
    const ofx_config &c = st.c;
    std::map<std::string, std::string> v;
    v["grid_step"] = std::to_string(c.grid_step);
    v["poc_win"] = std::to_string(c.poc_win);
    v["max_disp"] = py_float(c.max_disp);
    v["poc_peak_min"] = py_float(c.poc_peak_min);
    v["median_thresh"] = py_float(c.median_thresh);
    v["lam"] = py_float(c.lam);
    std::string sc = "[";
    for (int i = 0; i < c.n_scales; i++) sc += std::string(i ? "," : "") + "\n    " + py_float(c.scales[i]);
    v["scales"] = sc + "\n  ]";
    v["n_orient"] = std::to_string(c.n_orient);
    v["bandwidth"] = py_float(c.bandwidth);
    v["dc_cut"] = py_float(c.dc_cut);
    v["sigma_w"] = py_float(c.sigma_w);
    v["shape_order"] = std::to_string(c.shape_order);
    v["sigma_s"] = py_float(c.sigma_s);
    v["outer_iters"] = std::to_string(c.outer_iters);
    v["inner_iters"] = std::to_string(c.inner_iters);
    v["conf_min"] = py_float(c.conf_min);
    v["edge_guard"] = py_float(c.edge_guard);
    v["edge_guard_min_area"] = py_float(c.edge_guard_min_area);
    v["phase_tol"] = py_float(c.phase_tol);
    v["strain_win"] = std::to_string(c.strain_win);
    v["ref_update_conf"] = py_float(c.ref_update_conf);
    v["seq_mode"] = "\"" + st.seq_mode + "\"";
    v["fb_max"] = std::to_string(c.fb_max);
    v["fb_radius"] = py_float(c.fb_radius);
    v["fb_min_support"] = std::to_string(c.fb_min_support);
    v["fb_tol"] = py_float(c.fb_tol);
    v["tavg"] = std::to_string(st.tavg);
    v["tavg_align"] = st.tavg_align ? "true" : "false";
    v["tavg_model"] = "\"" + st.tavg_model + "\"";
    v["prewarp_freeze_at"] = std::to_string(c.prewarp_freeze_at);
    v["tile"] = std::to_string(c.tile);
    v["coarse_model"] = "\"" + st.coarse_model + "\"";
    v["coarse_gate"] = py_float(c.coarse_gate);
    std::string out = "{";
    bool first = true;
    for (const char *k : KEYS) {
        out += std::string(first ? "" : ",") + "\n  \"" + k + "\": " + v[k];
        first = false;
    }
    return out + "\n}";
}

void write_text(const std::string &path, const std::string &text) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(text.data(), 1, text.size(), f) != text.size()) {
        if (f) std::fclose(f);
        throw Fail("cannot write " + path);
    }
    std::fclose(f);
}

void make_dir(const std::string &d) {
    std::string part;
    for (size_t i = 0; i <= d.size(); i++) {
        if (i == d.size() || d[i] == '/') {
            part = d.substr(0, i);
            if (!part.empty() && ::mkdir(part.c_str(), 0777) != 0 && errno != EEXIST)
                throw Fail("cannot create " + part + ": " + std::strerror(errno));
        }
    }
}

std::string exe_dir() {
    char buf[PATH_MAX];
    const ssize_t n = ::readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    std::string s(buf);
    return s.substr(0, s.rfind('/'));
}

// the network of coarse_model "learned": weights/coarse_net.npy beside this program
void load_network(ofx_engine *e) {
    const std::string path = exe_dir() + "/weights/coarse_net.npy";
    const Pixels p = read_npy(path);
    if (p.bits != 32) throw Fail(path + ": expected float32 parameters");
    const size_t n = p.bytes.size() / sizeof(float);
    std::vector<float> v(n);
    std::memcpy(v.data(), p.bytes.data(), n * sizeof(float));
    check(ofx_coarse_net_load(e, v.data(), n));
}

// ------------------------------------------------------------------ command line
struct Args {
    std::string cmd, ref, def, frames, mask, out, config, crop = "auto", device = "gpu";
    std::vector<std::string> sets;
    long start = 0, stop = -1, step = 1, scale = 1, crop_margin = -1;
    bool raw = false;
};

long to_long(const std::string &opt, const std::string &v) {
    char *e;
    const long x = std::strtol(v.c_str(), &e, 10);
    if (v.empty() || *e) throw Fail("argument " + opt + ": invalid int value: '" + v + "'");
    return x;
}

Args parse(int argc, char **argv) {
    if (argc < 2 || (std::string(argv[1]) != "pair" && std::string(argv[1]) != "seq"))
        throw Fail("usage: ofx_cli {pair,seq} ... (see the source's header)");
    Args a;
    a.cmd = argv[1];
    for (int i = 2; i < argc; i++) {
        std::string opt = argv[i], val;
        const size_t eq = opt.find('=');
        const bool flag = opt == "--raw" || opt == "--png";
        if (!flag) {
            if (eq != std::string::npos) {
                val = opt.substr(eq + 1);
                opt = opt.substr(0, eq);
            } else {
                if (i + 1 >= argc) throw Fail("argument " + opt + ": expected one argument");
                val = argv[++i];
            }
        }
        if (opt == "--ref") a.ref = val;
        else if (opt == "--def") a.def = val;
        else if (opt == "--frames") a.frames = val;
        else if (opt == "--start") a.start = to_long(opt, val);
        else if (opt == "--stop") a.stop = to_long(opt, val);
        else if (opt == "--step") a.step = to_long(opt, val);
        else if (opt == "--mask") a.mask = val;
        else if (opt == "--scale") a.scale = to_long(opt, val);
        else if (opt == "--crop") a.crop = val;
        else if (opt == "--crop-margin") a.crop_margin = to_long(opt, val);
        else if (opt == "--config") a.config = val;
        else if (opt == "--set") a.sets.push_back(val);
        else if (opt == "--out") a.out = val;
        else if (opt == "--device") a.device = val;
        else if (opt == "--raw") a.raw = true;
        else if (opt == "--png") std::fprintf(stderr, "warning: --png is accepted for compatibility and writes nothing\n");
        else if (opt == "--roi") throw Fail("--roi polygons are not read by this twin: give the ROI as --mask");
        else throw Fail("unrecognized arguments: " + opt);
    }
    if (a.out.empty()) throw Fail("the following arguments are required: --out");
    if (a.cmd == "pair" && (a.ref.empty() || a.def.empty()))
        throw Fail("the following arguments are required: --ref, --def");
    if (a.cmd == "seq" && a.frames.empty()) throw Fail("the following arguments are required: --frames");
    if (a.scale < 1) throw Fail("scale must be >= 1, got " + std::to_string(a.scale));
    if (a.step < 1) throw Fail("step must be >= 1, got " + std::to_string(a.step));
    if (a.start < 0) throw Fail("start must be >= 0, got " + std::to_string(a.start));
    if (a.crop != "auto" && a.crop != "off") throw Fail("--crop must be auto or off, got " + a.crop);
    if (a.device != "gpu" && a.device != "hip") throw Fail("this build runs on the GPU only (--device gpu)");
    return a;
}

Settings settings(const Args &a) {
    Settings st;
    if (!a.config.empty()) {
        const std::string text = read_file(a.config);
        Parser p{text};
        for (auto &kv : p.object()) set_key(st, kv.first, kv.second);
    }
    for (const std::string &s : a.sets) {
        const size_t eq = s.find('=');
        if (eq == std::string::npos) throw Fail("--set needs KEY=VALUE, got '" + s + "'");
        std::string text = s.substr(eq + 1);
        // a bare word is a string value
        if (!text.empty() && (std::isalpha((unsigned char)text[0])) && text != "true" && text != "false")
            text = "\"" + text + "\"";
        Parser p{text};
        set_key(st, s.substr(0, eq), p.value());
    }
    if (st.tavg != 0) throw Fail("tavg is not run by this twin: average the frames with optflow_soil_amd.frames");
    return st;
}

struct Result {
    int w, h;
    std::vector<float> maps[10];
    std::vector<uint8_t> valid;
    std::vector<int32_t> fb;
    ofx_result r{};
    Result(int w_, int h_, bool with_fb) : w(w_), h(h_) {
        const size_t n = (size_t)w * h;
        for (auto &m : maps) m.assign(n, 0.f);
        valid.assign(n, 0);
        float **p[10] = {&r.ux, &r.uy, &r.conf, &r.raw_ux, &r.raw_uy, &r.exx, &r.eyy, &r.exy, &r.gmax, &r.vol};
        for (int k = 0; k < 10; k++) *p[k] = maps[k].data();
        r.valid = valid.data();
        if (with_fb) {
            fb.assign(n, 0);
            r.fb = fb.data();
        }
    }
};

std::vector<float> cut(const std::vector<float> &a, int w, const ofx_crop &c) {
    std::vector<float> out((size_t)(c.x1 - c.x0) * (c.y1 - c.y0));
    for (int y = c.y0; y < c.y1; y++)
        std::copy(a.begin() + (size_t)y * w + c.x0, a.begin() + (size_t)y * w + c.x1,
                  out.begin() + (size_t)(y - c.y0) * (c.x1 - c.x0));
    return out;
}
std::vector<uint8_t> cut(const std::vector<uint8_t> &a, int w, const ofx_crop &c) {
    std::vector<uint8_t> out((size_t)(c.x1 - c.x0) * (c.y1 - c.y0));
    for (int y = c.y0; y < c.y1; y++)
        std::copy(a.begin() + (size_t)y * w + c.x0, a.begin() + (size_t)y * w + c.x1,
                  out.begin() + (size_t)(y - c.y0) * (c.x1 - c.x0));
    return out;
}

int run(const Args &a) {
    Settings st = settings(a);
    // config.json first, as the reference's run does (run.rs:392-394)
    make_dir(a.out);
    write_text(a.out + "/config.json", config_json(st));
    // two lanes: an open sequence holds one, the frame conversion takes the other
    ofx_engine *e = nullptr;
    check(ofx_create(0, 2, &e));
    struct Close {
        ofx_engine *e;
        ~Close() { ofx_destroy(e); }
    } close{e};
    if (st.c.coarse_learned) load_network(e);
    std::vector<std::string> paths;
    if (a.cmd == "pair") {
        paths = {a.ref, a.def};
    } else {
        const std::vector<std::string> all = list_frames(a.frames);
        const long stop = a.stop < 0 ? (long)all.size() : std::min<long>(a.stop, (long)all.size());
        for (long k = a.start; k < stop; k += a.step) paths.push_back(all[k]);
        if (paths.size() < 2) throw Fail("a sequence needs at least two frames");
    }
    int w = 0, h = 0;
    std::vector<float> f0 = to_gray(e, read_pixels(paths[0]), (int)a.scale, w, h);
    std::vector<uint8_t> mask;
    if (!a.mask.empty()) mask = read_mask(a.mask, w, h);
    // the crop (run.rs:185-207): planned from the ROI, written as crop.json either way
    bool cropped = false;
    ofx_crop crop{};
    if (!mask.empty() && a.crop == "auto") {
        const int64_t m = a.crop_margin;
        check(ofx_crop_plan(mask.data(), w, h, &st.c, a.crop_margin >= 0 ? &m : nullptr, a.cmd == "seq", &crop));
        cropped = !(crop.x0 == 0 && crop.y0 == 0 && crop.x1 == w && crop.y1 == h);
        check(ofx_write_crop_json(a.out.c_str(), &crop));
    } else {
        check(ofx_write_crop_json(a.out.c_str(), nullptr));
    }
    const int cw = cropped ? crop.x1 - crop.x0 : w, ch = cropped ? crop.y1 - crop.y0 : h;
    auto view = [&](const std::vector<float> &f) { return cropped ? cut(f, w, crop) : f; };
    const std::vector<uint8_t> cmask = mask.empty() ? mask : (cropped ? cut(mask, w, crop) : mask);
    const ofx_crop *place = cropped ? &crop : nullptr;
    if (a.cmd == "pair") {
        int w2, h2;
        const std::vector<float> f1 = to_gray(e, read_pixels(paths[1]), (int)a.scale, w2, h2);
        if (w2 != w || h2 != h) throw Fail("images differ in size");
        Result res(cw, ch, false);
        check(ofx_track_pair(e, view(f0).data(), view(f1).data(), cw, ch, cmask.empty() ? nullptr : cmask.data(),
                             nullptr, nullptr, &st.c, &res.r));
        check(ofx_write_result(a.out.c_str(), &res.r, cw, ch, place, w, h, a.raw, 0));
        std::printf("pair %dx%d: coarse %.4f s, fine %.4f s, strain %.4f s\n", w, h, res.r.t_coarse, res.r.t_fine,
                    res.r.t_strain);
        return 0;
    }
    ofx_sequence *q = nullptr;
    check(ofx_seq_begin(e, view(f0).data(), cw, ch, cmask.empty() ? nullptr : cmask.data(), &st.c, &q));
    struct End {
        ofx_sequence *q;
        ~End() { ofx_seq_end(q); }
    } end{q};
    // Three stages overlap: a reader thread decodes frame k + 1 from its file while this thread
    // converts and tracks frame k on the GPU, and a writer thread writes frame k - 1's folder. Each
    // hand-off holds at most two items, which bounds the host memory at a few frames.
    const auto t_all = std::chrono::steady_clock::now();
    double t_read = 0, t_track = 0, t_write = 0;
    Channel<Pixels> decoded(2);
    struct Done {
        size_t k;
        std::unique_ptr<Result> res;
    };
    Channel<Done> tracked(2);
    std::thread reader([&] {
        try {
            for (size_t k = 1; k < paths.size(); k++) {
                const auto t0 = std::chrono::steady_clock::now();
                Pixels p = read_pixels(paths[k]);
                t_read += seconds_since(t0);
                if (!decoded.put(std::move(p))) return;
            }
            decoded.finish();
        } catch (...) {
            decoded.finish(std::current_exception());
        }
    });
    std::exception_ptr write_err;
    std::thread writer([&] {
        try {
            Done d;
            while (tracked.take(d)) {
                const auto t0 = std::chrono::steady_clock::now();
                char dir[32];
                std::snprintf(dir, sizeof dir, "/frame_%05zu", d.k);
                check(ofx_write_result((a.out + dir).c_str(), &d.res->r, cw, ch, place, w, h, a.raw, 1));
                t_write += seconds_since(t0);
                std::printf("frame %zu: reference %d, mean confidence %.4f\n", d.k, d.res->r.reference,
                            d.res->r.mean_conf);
            }
        } catch (...) {
            write_err = std::current_exception();
            tracked.abandon();
        }
    });
    // stops both threads on every way out of this scope, an error here included
    struct Join {
        Channel<Pixels> &decoded;
        Channel<Done> &tracked;
        std::thread &reader, &writer;
        ~Join() {
            decoded.abandon();
            tracked.finish();
            if (reader.joinable()) reader.join();
            if (writer.joinable()) writer.join();
        }
    } join{decoded, tracked, reader, writer};
    Pixels px;
    for (size_t k = 1; decoded.take(px); k++) {
        const auto t0 = std::chrono::steady_clock::now();
        int wk, hk;
        const std::vector<float> fk = to_gray(e, px, (int)a.scale, wk, hk);
        if (wk != w || hk != h) throw Fail("frame " + paths[k] + " differs in size from the first");
        auto res = std::make_unique<Result>(cw, ch, st.c.seq_hybrid != 0);
        check(ofx_seq_step(q, view(fk).data(), &res->r));
        t_track += seconds_since(t0);
        if (!tracked.put(Done{k, std::move(res)})) break;
    }
    tracked.finish();
    writer.join();
    if (write_err) std::rethrow_exception(write_err);
    reader.join();
    std::fprintf(stderr, "seq: %zu frames in %.3f s; read %.3f s, track %.3f s, write %.3f s, overlapped\n",
                 paths.size() - 1, seconds_since(t_all), t_read, t_track, t_write);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        return run(parse(argc, argv));
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 2;
    }
}
