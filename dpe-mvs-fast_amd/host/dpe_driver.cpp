// dpe_driver.cpp -- the C++ host driver: main.cpp's CLI, problem list,
// multi-scale schedule, per-image process calls, state hand-off, checkpoints,
// fusion and profile summary (main.cpp:1-466), over libdpe_hip.so (GPU) and
// libdpe_host.so (formats, pyramid, support maps).
//
//   dpe_driver input_folder [gpu_index] [--output=folder] [--vis=none|final|all]
//              [--checkpoint=none|final|all] [--profile=on|off]
//              [--schedule=file] [--support=own|reference]
//              [--fusion=eth|tat_intermediate|tat_advanced] [--dry-run]
//
// Input folder: images/%08d.jpg (sequential 8-bit JPEG, decoded as libjpeg
// does for cv::imread), .pgm (8-bit gray) or .ppm (8-bit colour), cams/%08d_cam.txt,
// pair.txt (main.cpp:336-341).  Output: <output>/views/%08d/{depths.dmb,
// normals.dmb, weak.bin, selected_views.bin, edges_<s>.dmb, labels_<s>.dmb}
// and <output>/DPE.ply.
//
// --schedule=file replaces the reference schedule (main.cpp:390-457) with one
// call group per line (applied to every problem, in order):
//   call scale=S state=FIRST_INIT|REFINE_INIT|REFINE_ITER iters=N apd=0|1 edge=0|1 geom=0|1 peak=P
//        [ransac=F] [rotate=R]
//   fusion min_consistent=K
// --support=reference computes the edge and label maps with the reference's
// EdgeSegment pipeline (Canny edges, Roberts + Hough + Connect labels;
// dpe_host_edge_segment), --support=own (default) with dpe_host_support_maps.
// --fusion selects RunFusion (eth, default; DPE.cpp:1387-1548) or one of the
// k-of-n Tanks-and-Temples fusions (RunFusion_TAT_Intermediate / _advanced,
// DPE.cpp:1550-1867), which read blocks/mask_<id>.jpg when that folder exists.
// --dry-run prints the problems and every call of the schedule, then exits
// without touching a GPU.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "dpe_hip.h"
#include "dpe_host.h"

namespace {

enum class Mode { NONE, FINAL, ALL };

struct Options {
    int gpu = 0;
    Mode vis = Mode::NONE;
    Mode checkpoint = Mode::FINAL;
    bool profile = true;
    bool dry_run = false;
    bool reference_support = false;
    int fusion = -1;  // -1 RunFusion, else DPE_TAT_INTERMEDIATE / DPE_TAT_ADVANCED
    std::string output, schedule;
};

struct CallSpec {
    int scale = 1;
    int state = DPE_FIRST_INIT;
    int iters = 3;
    int apd = 0, edge = 0, geom = 0, peak = 6;
    float ransac = 0.005f;  // PatchMatchParams default (main.h:81-109)
    int rotate = 4;
    bool final_call = false;  // the call whose results count as "final" for checkpoints
};

struct Problem {
    int index = 0;
    int ref = 0;
    std::vector<int> srcs;
    std::string result_folder;
};

struct ViewState {
    int w = 0, h = 0;
    std::vector<float> depth, normal;  // normal: 3 floats per pixel (world frame)
    std::vector<uint8_t> weak;
    std::vector<uint32_t> selected;
};

struct Profile {
    double edge_ms = 0, input_ms = 0, support_ms = 0, gpu_call_ms = 0, gpu_kernel_ms = 0, result_ms = 0,
           checkpoint_ms = 0, fusion_ms = 0;
    int process_calls = 0;
};

double ms_since(const std::chrono::steady_clock::time_point &t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

std::string fmt8(int i) {
    std::ostringstream ss;
    ss << std::setw(8) << std::setfill('0') << i;
    return ss.str();
}

bool exists(const std::string &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

void make_dirs(const std::string &p) {
    std::string cur;
    std::istringstream ss(p);
    std::string part;
    if (!p.empty() && p[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
}

// ---------------------------------------------------------------- images
// Binary PGM (P5) or PPM (P6), maxval 255.  Colour is returned as BGR.
bool read_pnm(const std::string &path, int &w, int &h, int &channels, std::vector<uint8_t> &data) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::string magic;
    in >> magic;
    if (magic != "P5" && magic != "P6") return false;
    channels = magic == "P5" ? 1 : 3;
    auto next_int = [&](int &v) {
        for (;;) {
            in >> std::ws;
            if (in.peek() == '#') {
                std::string skip;
                std::getline(in, skip);
                continue;
            }
            in >> v;
            return (bool)in;
        }
    };
    int maxval = 0;
    if (!next_int(w) || !next_int(h) || !next_int(maxval) || maxval != 255) return false;
    in.get();
    data.resize((size_t)w * h * channels);
    in.read((char *)data.data(), (std::streamsize)data.size());
    if (!in) return false;
    if (channels == 3)
        for (size_t i = 0; i < (size_t)w * h; ++i) std::swap(data[3 * i], data[3 * i + 2]);  // RGB -> BGR
    return true;
}

// BGR -> gray with the BT.601 integer weights cv::cvtColor uses (14-bit fixed point).
uint8_t bgr_to_gray(const uint8_t *p) {
    return (uint8_t)((p[0] * 1868 + p[1] * 9617 + p[2] * 4899 + (1 << 13)) >> 14);
}

struct SceneCache {
    std::string image_dir, cam_dir;
    std::map<int, std::vector<uint8_t>> gray;
    std::map<int, std::pair<int, int>> size;
    std::map<int, dpe_camera> cams;
    std::map<std::pair<int, int>, std::vector<float>> float_images;  // (id, scale)

    std::string image_path(int id, const char *ext) const { return image_dir + "/" + fmt8(id) + ext; }

    const std::vector<uint8_t> &gray_image(int id) {
        auto it = gray.find(id);
        if (it != gray.end()) return it->second;
        int w = 0, h = 0, ch = 0;
        std::vector<uint8_t> data;
        // images/%08d.jpg as cv::imread(IMREAD_GRAYSCALE) reads it (DPE.cpp:713), else PGM / PPM
        const std::string jpg = image_path(id, ".jpg");
        if (exists(jpg)) {
            const int st = dpe_host_read_jpeg(jpg.c_str(), 0, nullptr, 0, &w, &h);
            data.resize((size_t)w * h);
            if (st != 0 || dpe_host_read_jpeg(jpg.c_str(), 0, data.data(), data.size(), &w, &h) != 0)
                throw std::runtime_error("Cannot decode image (sequential 8-bit JPEG only): " + jpg);
            ch = 1;
        } else if (!read_pnm(image_path(id, ".pgm"), w, h, ch, data) &&
                   !read_pnm(image_path(id, ".ppm"), w, h, ch, data)) {
            throw std::runtime_error("Cannot read image: " + image_path(id, ".jpg") + " / .pgm / .ppm");
        }
        std::vector<uint8_t> g((size_t)w * h);
        if (ch == 1) g = data;
        else
            for (size_t i = 0; i < g.size(); ++i) g[i] = bgr_to_gray(&data[3 * i]);
        size[id] = {w, h};
        return gray.emplace(id, std::move(g)).first->second;
    }

    // BGR colour image at full resolution (gray replicated for PGM input)
    std::vector<uint8_t> color_image(int id, int &w, int &h) {
        int ch = 0;
        std::vector<uint8_t> data;
        // cv::imread(IMREAD_COLOR) (DPE.cpp:1422)
        const std::string jpg = image_path(id, ".jpg");
        if (exists(jpg) && dpe_host_read_jpeg(jpg.c_str(), 1, nullptr, 0, &w, &h) == 0) {
            data.resize((size_t)w * h * 3);
            if (dpe_host_read_jpeg(jpg.c_str(), 1, data.data(), data.size(), &w, &h) == 0) return data;
        }
        if (read_pnm(image_path(id, ".ppm"), w, h, ch, data) && ch == 3) return data;
        const std::vector<uint8_t> &g = gray_image(id);
        w = size[id].first;
        h = size[id].second;
        std::vector<uint8_t> c(g.size() * 3);
        for (size_t i = 0; i < g.size(); ++i) c[3 * i] = c[3 * i + 1] = c[3 * i + 2] = g[i];
        return c;
    }

    // GetGrayFloatImage (DPE.cpp:720-737): size round(W / s) x round(H / s)
    const std::vector<float> &float_image(int id, int scale, int &w, int &h) {
        const std::vector<uint8_t> &g = gray_image(id);
        const int W = size[id].first, H = size[id].second;
        w = (int)std::lround(W * (1.0f / scale));
        h = (int)std::lround(H * (1.0f / scale));
        auto key = std::make_pair(id, scale);
        auto it = float_images.find(key);
        if (it != float_images.end()) return it->second;
        std::vector<float> full(g.begin(), g.end());
        std::vector<float> out((size_t)w * h);
        dpe_host_resize_linear(full.data(), W, H, out.data(), w, h);
        return float_images.emplace(key, std::move(out)).first->second;
    }

    // GetCamera (DPE.cpp:739-760): K scaled to the target size
    dpe_camera camera(int id, int tw, int th) {
        auto it = cams.find(id);
        if (it == cams.end()) {
            dpe_camera c{};
            const std::string p = cam_dir + "/" + fmt8(id) + "_cam.txt";
            if (dpe_host_read_camera(p.c_str(), &c) != 0) throw std::runtime_error("Cannot read camera: " + p);
            it = cams.emplace(id, c).first;
        }
        dpe_camera c = it->second;
        gray_image(id);
        const float sx = tw / (float)size[id].first, sy = th / (float)size[id].second;
        c.K[0] *= sx;
        c.K[2] *= sx;
        c.K[4] *= sy;
        c.K[5] *= sy;
        c.width = tw;
        c.height = th;
        return c;
    }
};

template <class T>
std::vector<T> rescaled(const std::vector<T> &src, int w, int h, int dw, int dh, int elems = 1) {
    std::vector<T> out((size_t)dw * dh * elems);
    dpe_host_rescale_nearest(src.data(), w, h, out.data(), dw, dh, (int)(sizeof(T) * elems));
    return out;
}

// ---------------------------------------------------------------- schedule
std::vector<CallSpec> reference_schedule(int round_num) {
    // main.cpp:390-457: per round one INIT call, then 3 REFINE_ITER calls
    std::vector<CallSpec> calls;
    for (int i = 0; i < round_num; ++i) {
        const int scale = 1 << (round_num - 1 - i);
        CallSpec c;
        c.scale = scale;
        c.state = i == 0 ? DPE_FIRST_INIT : DPE_REFINE_INIT;
        c.apd = c.edge = i == 0 ? 0 : 1;
        c.ransac = i == 0 ? 0.005f : (float)(0.01 - i * 0.00125);  // round 0 keeps the default
        c.rotate = i == 0 ? 4 : std::min(1 << i, 4);
        c.geom = 0;
        c.iters = 3;
        c.peak = 6;
        calls.push_back(c);
        for (int j = 0; j < 3; ++j) {
            CallSpec r;
            r.scale = scale;
            r.state = DPE_REFINE_ITER;
            r.apd = r.edge = i == 0 ? 0 : 1;
            r.ransac = (float)(0.01 - i * 0.00125);
            r.rotate = std::min(1 << i, 4);
            r.geom = 1;
            r.iters = 3;
            r.peak = std::max(4 - 2 * j, 2);
            r.final_call = i == round_num - 1 && j == 2;
            calls.push_back(r);
        }
    }
    return calls;
}

std::vector<CallSpec> read_schedule(const std::string &path, int &min_consistent) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot read schedule: " + path);
    std::vector<CallSpec> calls;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string kind;
        if (!(ss >> kind) || kind[0] == '#') continue;
        std::map<std::string, std::string> kv;
        std::string tok;
        while (ss >> tok) {
            const size_t eq = tok.find('=');
            if (eq == std::string::npos) throw std::runtime_error("bad schedule token: " + tok);
            kv[tok.substr(0, eq)] = tok.substr(eq + 1);
        }
        if (kind == "fusion") {
            if (kv.count("min_consistent")) min_consistent = std::atoi(kv["min_consistent"].c_str());
            continue;
        }
        if (kind != "call") throw std::runtime_error("bad schedule line: " + line);
        CallSpec c;
        if (kv.count("scale")) c.scale = std::atoi(kv["scale"].c_str());
        if (kv.count("state")) {
            const std::string s = kv["state"];
            c.state = s == "FIRST_INIT" ? DPE_FIRST_INIT : s == "REFINE_INIT" ? DPE_REFINE_INIT
                      : s == "REFINE_ITER" ? DPE_REFINE_ITER : -1;
            if (c.state < 0) throw std::runtime_error("bad state: " + s);
        }
        if (kv.count("iters")) c.iters = std::atoi(kv["iters"].c_str());
        if (kv.count("apd")) c.apd = std::atoi(kv["apd"].c_str());
        if (kv.count("edge")) c.edge = std::atoi(kv["edge"].c_str());
        if (kv.count("geom")) c.geom = std::atoi(kv["geom"].c_str());
        if (kv.count("peak")) c.peak = std::atoi(kv["peak"].c_str());
        if (kv.count("ransac")) c.ransac = std::strtof(kv["ransac"].c_str(), nullptr);
        if (kv.count("rotate")) c.rotate = std::atoi(kv["rotate"].c_str());
        if (c.scale < 1 || (c.scale & (c.scale - 1))) throw std::runtime_error("scale must be a power of 2");
        calls.push_back(c);
    }
    if (calls.empty()) throw std::runtime_error("schedule has no calls: " + path);
    calls.back().final_call = true;
    return calls;
}

int log2i(int v) {
    int s = 0;
    while ((1 << s) < v) s++;
    return s;
}

const char *state_name(int s) {
    return s == DPE_FIRST_INIT ? "FIRST_INIT" : s == DPE_REFINE_INIT ? "REFINE_INIT" : "REFINE_ITER";
}

// ---------------------------------------------------------------- driver
struct Driver {
    Options opt;
    std::string dense, output;
    SceneCache cache;
    std::vector<Problem> problems;
    std::unordered_map<int, ViewState> store;
    Profile prof;
    dpe_ctx *ctx = nullptr;
    int max_scale_size = 1;

    void check(dpe_status st, const char *what) {
        if (st != DPE_OK)
            throw std::runtime_error(std::string(what) + ": " + dpe_last_error(ctx));
    }

    void load_problems() {
        const std::string pair = dense + "/pair.txt";
        std::vector<int32_t> refs(4096), counts(4096), srcs((size_t)4096 * DPE_MAX_IMAGES);
        const int n = dpe_host_read_pairs(pair.c_str(), 4096, refs.data(), counts.data(), srcs.data());
        if (n < 0) throw std::runtime_error("Cannot read " + pair);
        for (int i = 0; i < n && i < 4096; ++i) {
            Problem p;
            p.index = i;
            p.ref = refs[i];
            for (int j = 0; j < std::min<int>(counts[i], DPE_MAX_IMAGES); ++j)
                p.srcs.push_back(srcs[(size_t)i * DPE_MAX_IMAGES + j]);
            if (counts[i] + 1 > DPE_MAX_IMAGES)  // InuputInitialization's limit (DPE.cpp:944-947)
                throw std::runtime_error("Can't process so much images: " + std::to_string(counts[i] + 1));
            p.result_folder = output + "/views/" + fmt8(p.ref);
            problems.push_back(p);
        }
    }

    // CheckImages (main.cpp:147-164): all reference images share one size
    void check_images() {
        if (problems.empty()) throw std::runtime_error("Images may error, check it!");
        cache.gray_image(problems[0].ref);
        const auto sz = cache.size[problems[0].ref];
        for (const Problem &p : problems) {
            cache.gray_image(p.ref);
            if (cache.size[p.ref] != sz) throw std::runtime_error("Images may error, check it!");
        }
    }

    // ComputeRoundNum (main.cpp:223-238)
    int round_num() {
        const auto sz = cache.size[problems[0].ref];
        int m = std::max(sz.first, sz.second), r = 1;
        while (m > 800) {
            m /= 2;
            r++;
        }
        return r;
    }

    // GetProblemEdges (main.cpp:166-221): support maps of one problem at one
    // scale, cached as .dmb in the result folder
    void problem_edges(const Problem &p, int scale_size) {
        const int s = log2i(scale_size);
        const std::string ep = p.result_folder + "/edges_" + std::to_string(s) + ".dmb";
        const std::string lp = p.result_folder + "/labels_" + std::to_string(s) + ".dmb";
        if (exists(ep) && exists(lp)) return;
        int w = 0, h = 0;
        const std::vector<float> &f = cache.float_image(p.ref, scale_size, w, h);
        std::vector<uint8_t> u8(f.size());
        for (size_t i = 0; i < f.size(); ++i) {  // convertTo(CV_8UC1): round half to even, saturate
            const float v = std::nearbyint(f[i]);
            u8[i] = (uint8_t)std::min(255.0f, std::max(0.0f, v));
        }
        std::vector<uint8_t> edges(u8.size());
        std::vector<int32_t> labels(u8.size());
        if (opt.reference_support) {
            // GetProblemEdges (main.cpp:166-221): edges by Canny on the scaled image,
            // labels from the full-size image reduced to the call's size
            int fw = 0, fh = 0, ow = 0, oh = 0;
            const std::vector<float> &full = cache.float_image(p.ref, 1, fw, fh);
            std::vector<uint8_t> full8(full.size());
            for (size_t i = 0; i < full.size(); ++i)
                full8[i] = (uint8_t)std::min(255.0f, std::max(0.0f, std::nearbyint(full[i])));
            if (dpe_host_edge_segment(s, u8.data(), w, h, 0, 1, 1, edges.data(), nullptr, &ow, &oh) != 0 ||
                ow != w || oh != h)
                throw std::runtime_error("EdgeSegment (edges) failed for " + p.result_folder);
            std::vector<int32_t> lab((size_t)(std::lround(fw / (double)(1 << s)) + 1) *
                                     (std::lround(fh / (double)(1 << s)) + 1));
            if (dpe_host_edge_segment(s, full8.data(), fw, fh, 1, 0, 1, nullptr, lab.data(), &ow, &oh) != 0 ||
                ow != w || oh != h)
                throw std::runtime_error("EdgeSegment (labels) does not match the call's size for " +
                                         p.result_folder);
            std::copy(lab.begin(), lab.begin() + (size_t)w * h, labels.begin());
        } else {
            // labels at the call's size with 1/1024 of its area as the small-region limit
            dpe_host_support_maps(u8.data(), w, h, 0, edges.data(), labels.data());
        }
        if (dpe_host_write_bin_mat(ep.c_str(), h, w, DPE_HOST_CV_8UC1, edges.data()) != 0 ||
            dpe_host_write_bin_mat(lp.c_str(), h, w, DPE_HOST_CV_32SC1, labels.data()) != 0)
            throw std::runtime_error("Cannot write support maps to " + p.result_folder);
    }

    template <class T>
    std::vector<T> read_mat(const std::string &path, int expect_type, int &w, int &h) {
        int rows = 0, cols = 0, type = 0;
        if (dpe_host_read_bin_mat(path.c_str(), &rows, &cols, &type, nullptr, 0) != 0 || type != expect_type)
            throw std::runtime_error("Cannot read " + path);
        std::vector<T> v((size_t)rows * cols * dpe_host_type_size(type) / sizeof(T));
        if (dpe_host_read_bin_mat(path.c_str(), &rows, &cols, &type, v.data(), v.size() * sizeof(T)) != 0)
            throw std::runtime_error("Cannot read " + path);
        w = cols;
        h = rows;
        return v;
    }

    // ProcessProblem (main.cpp:241-319) for one call of one problem
    void process(const Problem &p, const CallSpec &c, int iteration, bool write_checkpoint) {
        auto t0 = std::chrono::steady_clock::now();
        std::cout << "Processing image: " << fmt8(p.ref) << "..." << std::endl;
        // ---- InuputInitialization (DPE.cpp:934-1069)
        int W = 0, H = 0;
        std::vector<const std::vector<float> *> imgs;
        imgs.push_back(&cache.float_image(p.ref, c.scale, W, H));
        for (int s : p.srcs) {
            int w = 0, h = 0;
            imgs.push_back(&cache.float_image(s, c.scale, w, h));
            if (w != W || h != H) throw std::runtime_error("source image size differs from the reference");
        }
        dpe_params prm{};
        prm.max_iterations = c.iters;
        prm.sigma_spatial = 5.0f;
        prm.sigma_color = 3.0f;
        prm.top_k = 4;
        prm.geom_consistency = c.geom;
        prm.strong_radius = 5;
        prm.strong_increment = 2;
        prm.weak_radius = 5;
        prm.weak_increment = 5;
        prm.use_APD = c.apd;
        prm.use_edge = c.edge;
        prm.use_limit = 1;
        prm.use_label = 1;
        prm.use_radius = 1;
        prm.high_res_img = 1;
        prm.max_scale_size = max_scale_size;
        prm.scale_size = c.scale;
        prm.weak_peak_radius = c.peak;
        prm.rotate_time = c.rotate;
        prm.ransac_threshold = c.ransac;
        prm.geom_factor = 0.2f;
        prm.state = c.state;
        prm.call_index = iteration;
        prm.seed = 0x5EED0002ull;
        dpe_inputs in{};
        in.width = W;
        in.height = H;
        in.num_images = (int)imgs.size();
        in.memory = DPE_MEM_HOST;
        std::vector<int> ids{p.ref};
        ids.insert(ids.end(), p.srcs.begin(), p.srcs.end());
        for (size_t i = 0; i < ids.size(); ++i) {
            in.images[i] = imgs[i]->data();
            in.cameras[i] = cache.camera(ids[i], W, H);
        }
        prm.depth_min = in.cameras[0].depth_min * 0.6f;
        prm.depth_max = in.cameras[0].depth_max * 1.2f;
        prm.num_images = in.num_images;
        std::vector<std::vector<float>> geom(ids.size());
        if (c.geom) {
            for (size_t i = 0; i < ids.size(); ++i) {
                auto it = store.find(ids[i]);
                if (it == store.end()) geom[i].assign((size_t)W * H, 0.0f);
                else geom[i] = rescaled(it->second.depth, it->second.w, it->second.h, W, H);
                in.geom_depths[i] = geom[i].data();
            }
        }
        std::vector<uint8_t> weak;
        std::vector<float> prior;
        std::vector<uint32_t> prior_sel;
        const auto st_it = store.find(p.ref);
        if (c.apd) {
            if (st_it == store.end()) throw std::runtime_error("Can't find weak info of " + fmt8(p.ref));
            weak = rescaled(st_it->second.weak, st_it->second.w, st_it->second.h, W, H);
            in.weak_info = weak.data();
        }
        if (c.state != DPE_FIRST_INIT) {
            if (st_it == store.end()) throw std::runtime_error("no prior state of " + fmt8(p.ref));
            const ViewState &vs = st_it->second;
            const std::vector<float> d = rescaled(vs.depth, vs.w, vs.h, W, H);
            const std::vector<float> n = rescaled(vs.normal, vs.w, vs.h, W, H, 3);
            prior_sel = rescaled(vs.selected, vs.w, vs.h, W, H);
            prior.resize((size_t)4 * W * H);
            for (size_t i = 0; i < (size_t)W * H; ++i) {
                prior[4 * i] = n[3 * i];
                prior[4 * i + 1] = n[3 * i + 1];
                prior[4 * i + 2] = n[3 * i + 2];
                prior[4 * i + 3] = d[i];
            }
            in.prior_planes = prior.data();
            in.prior_selected = prior_sel.data();
        }
        prof.input_ms += ms_since(t0);
        // ---- SupportInitialization (DPE.cpp:1187-1214)
        auto t1 = std::chrono::steady_clock::now();
        const int s = log2i(c.scale), smax = log2i(max_scale_size);
        int ew = 0, eh = 0, lw = 0, lh = 0, ow = 0, oh = 0;
        std::vector<uint8_t> edges =
            read_mat<uint8_t>(p.result_folder + "/edges_" + std::to_string(s) + ".dmb", DPE_HOST_CV_8UC1, ew, eh);
        std::vector<uint8_t> edges_low =
            read_mat<uint8_t>(p.result_folder + "/edges_" + std::to_string(smax) + ".dmb", DPE_HOST_CV_8UC1, lw, lh);
        std::vector<int32_t> labels =
            read_mat<int32_t>(p.result_folder + "/labels_" + std::to_string(s) + ".dmb", DPE_HOST_CV_32SC1, ow, oh);
        if (ew != W || eh != H || ow != W || oh != H) throw std::runtime_error("support maps do not match the image");
        in.edges = edges.data();
        in.edges_low = edges_low.data();
        in.low_width = lw;
        in.low_height = lh;
        in.labels = labels.data();
        prof.support_ms += ms_since(t1);
        // ---- RunPatchMatch
        auto t2 = std::chrono::steady_clock::now();
        std::vector<float> planes((size_t)4 * W * H), costs((size_t)W * H);
        std::vector<uint8_t> states((size_t)W * H);
        std::vector<uint32_t> selected((size_t)W * H);
        dpe_outputs out{planes.data(), costs.data(), states.data(), selected.data()};
        check(dpe_process(ctx, &prm, &in, &out), "dpe_process");
        dpe_stats st{};
        check(dpe_last_stats(ctx, &st), "dpe_last_stats");
        prof.gpu_call_ms += ms_since(t2);
        prof.gpu_kernel_ms += st.gpu_ms;
        // ---- result assembly (main.cpp:264-280)
        auto t3 = std::chrono::steady_clock::now();
        ViewState vs;
        vs.w = W;
        vs.h = H;
        vs.depth.resize((size_t)W * H);
        vs.normal.resize((size_t)3 * W * H);
        for (size_t i = 0; i < (size_t)W * H; ++i) {
            float d = planes[4 * i + 3];
            if (d < prm.depth_min || d > prm.depth_max) {
                d = 0.0f;
                states[i] = DPE_UNKNOWN;
            }
            vs.depth[i] = d;
            vs.normal[3 * i] = planes[4 * i];
            vs.normal[3 * i + 1] = planes[4 * i + 1];
            vs.normal[3 * i + 2] = planes[4 * i + 2];
        }
        vs.weak = std::move(states);
        vs.selected = std::move(selected);
        store[p.ref] = std::move(vs);
        prof.result_ms += ms_since(t3);
        if (write_checkpoint) {
            auto t4 = std::chrono::steady_clock::now();
            const ViewState &v = store[p.ref];
            const std::string f = p.result_folder;
            bool ok = dpe_host_write_bin_mat((f + "/depths.dmb").c_str(), H, W, DPE_HOST_CV_32FC1, v.depth.data()) == 0;
            ok = ok && dpe_host_write_bin_mat((f + "/normals.dmb").c_str(), H, W, DPE_HOST_CV_32FC3, v.normal.data()) == 0;
            ok = ok && dpe_host_write_bin_mat((f + "/weak.bin").c_str(), H, W, DPE_HOST_CV_8UC1, v.weak.data()) == 0;
            ok = ok && dpe_host_write_bin_mat((f + "/selected_views.bin").c_str(), H, W, DPE_HOST_CV_32SC1,
                                              v.selected.data()) == 0;
            if (!ok) throw std::runtime_error("Cannot write checkpoint to " + f);
            prof.checkpoint_ms += ms_since(t4);
        }
        prof.process_calls++;
        std::cout << "Processing image: " << fmt8(p.ref) << " done! (" << W << "x" << H << ", " << state_name(c.state)
                  << ", GPU " << st.gpu_ms << " ms)" << std::endl;
    }

    // RunFusion (DPE.cpp:1387-1548) through dpe_fuse, or a TAT fusion
    // (DPE.cpp:1550-1867) through dpe_fuse_tat, then ExportPointCloud
    int64_t fuse(int min_consistent) {
        std::unordered_map<int, int> index;
        for (const Problem &p : problems) index.emplace(p.ref, p.index);
        const int n = (int)problems.size();
        std::vector<dpe_fuse_view> views(n);
        std::vector<std::vector<uint8_t>> colors(n);
        size_t total = 0;
        for (int i = 0; i < n; ++i) {
            const Problem &p = problems[i];
            const ViewState &vs = store.at(p.ref);
            int cw = 0, ch = 0;
            std::vector<uint8_t> color = cache.color_image(p.ref, cw, ch);
            dpe_camera cam{};
            const std::string cp = cache.cam_dir + "/" + fmt8(p.ref) + "_cam.txt";
            if (dpe_host_read_camera(cp.c_str(), &cam) != 0) throw std::runtime_error("Cannot read camera: " + cp);
            cam.width = cw;
            cam.height = ch;
            if (cw != vs.w || ch != vs.h) {
                // RescaleImageAndCamera (DPE.cpp:1290-1311): bilinear colour, scaled K
                std::vector<uint8_t> scaled((size_t)vs.w * vs.h * 3);
                for (int k = 0; k < 3; ++k) {
                    std::vector<float> plane((size_t)cw * ch), out((size_t)vs.w * vs.h);
                    for (size_t j = 0; j < plane.size(); ++j) plane[j] = color[3 * j + k];
                    dpe_host_resize_linear(plane.data(), cw, ch, out.data(), vs.w, vs.h);
                    for (size_t j = 0; j < out.size(); ++j)
                        scaled[3 * j + k] = (uint8_t)std::min(255.0f, std::max(0.0f, std::nearbyint(out[j])));
                }
                color.swap(scaled);
                const float sx = vs.w / (float)cw, sy = vs.h / (float)ch;
                cam.K[0] *= sx;
                cam.K[2] *= sx;
                cam.K[4] *= sy;
                cam.K[5] *= sy;
                cam.width = vs.w;
                cam.height = vs.h;
            }
            colors[i] = std::move(color);
            dpe_fuse_view &v = views[i];
            v.width = vs.w;
            v.height = vs.h;
            v.camera = cam;
            v.depth = vs.depth.data();
            v.normal = vs.normal.data();
            v.pixel_states = vs.weak.data();
            v.color = colors[i].data();
            v.num_src = (int)p.srcs.size();
            for (size_t j = 0; j < p.srcs.size(); ++j) {
                // imageIdToindexMap[id] (DPE.cpp:1457): an id without a problem maps to 0
                auto it = index.find(p.srcs[j]);
                v.src_index[j] = it == index.end() ? 0 : it->second;
            }
            total += (size_t)vs.w * vs.h;
        }
        std::vector<dpe_point> points(total);
        int64_t count = 0;
        if (opt.fusion < 0) {
            dpe_fuse_params fp{};
            fp.max_reproj_px = 2.0f;
            fp.max_rel_depth = 0.01f;
            fp.max_angle_rad = 0.174533f;
            fp.strong_factor = 0.3f;
            fp.weak_factor = 0.45f;
            fp.min_consistent = min_consistent;
            fp.memory = DPE_MEM_HOST;
            check(dpe_fuse(ctx, views.data(), n, &fp, points.data(), (int64_t)points.size(), &count), "dpe_fuse");
        } else {
            // blocks/mask_<id>.jpg, read as cv::imread(IMREAD_GRAYSCALE) and indexed
            // at the depth map's (r, c) (DPE.cpp:1574-1600, 1643)
            std::vector<std::vector<uint8_t>> blocks;
            std::vector<const uint8_t *> block_ptrs;
            if (exists(dense + "/blocks")) {
                for (int i = 0; i < n; ++i) {
                    const std::string bp = dense + "/blocks/mask_" + std::to_string(problems[i].ref) + ".jpg";
                    int bw = 0, bh = 0;
                    std::vector<uint8_t> img;
                    if (dpe_host_read_jpeg(bp.c_str(), 0, nullptr, 0, &bw, &bh) == 0) {
                        img.resize((size_t)bw * bh);
                        if (dpe_host_read_jpeg(bp.c_str(), 0, img.data(), img.size(), &bw, &bh) != 0) img.clear();
                    }
                    if (img.empty()) throw std::runtime_error("Cannot read block mask: " + bp);
                    if (bw < views[i].width || bh < views[i].height)
                        throw std::runtime_error("Block mask smaller than the depth map: " + bp);
                    std::vector<uint8_t> b((size_t)views[i].width * views[i].height);
                    for (int r = 0; r < views[i].height; ++r)
                        std::memcpy(&b[(size_t)r * views[i].width], &img[(size_t)r * bw], views[i].width);
                    blocks.push_back(std::move(b));
                }
                for (const auto &b : blocks) block_ptrs.push_back(b.data());
            }
            dpe_fuse_tat_params tp{};
            tp.variant = opt.fusion;
            tp.dist_base = 0.25f;
            tp.depth_base = opt.fusion == DPE_TAT_INTERMEDIATE ? 1.0f / 3500.0f : 1.0f / 3000.0f;
            tp.angle_base = 0.06981317007977318f;
            tp.angle_grad = 0.05235987755982988f;
            tp.memory = DPE_MEM_HOST;
            check(dpe_fuse_tat(ctx, views.data(), n, block_ptrs.empty() ? nullptr : block_ptrs.data(), &tp,
                               points.data(), (int64_t)points.size(), &count),
                  "dpe_fuse_tat");
        }
        const std::string ply = output + "/DPE.ply";
        if (dpe_host_write_ply(ply.c_str(), points.data(), count) != 0) throw std::runtime_error("Cannot write " + ply);
        std::cout << "[DPE Progress] fusion done; point cloud saved to " << ply << " with " << count << " points"
                  << std::endl;
        return count;
    }
};

Options parse_options(int argc, char **argv) {
    Options o;
    bool gpu_set = false;
    auto mode = [](const std::string &v, const char *what) {
        if (v == "none") return Mode::NONE;
        if (v == "final") return Mode::FINAL;
        if (v == "all") return Mode::ALL;
        throw std::runtime_error(std::string("Invalid ") + what + " value: " + v);
    };
    for (int i = 2; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a.rfind("--vis=", 0) == 0) o.vis = mode(a.substr(6), "--vis");
        else if (a.rfind("--checkpoint=", 0) == 0) o.checkpoint = mode(a.substr(13), "--checkpoint");
        else if (a.rfind("--profile=", 0) == 0) o.profile = a.substr(10) != "off" && a.substr(10) != "0";
        else if (a.rfind("--output=", 0) == 0) {
            o.output = a.substr(9);
            if (o.output.empty()) throw std::runtime_error("--output requires a non-empty path");
        } else if (a == "--output") {
            if (i + 1 >= argc) throw std::runtime_error("--output requires a path");
            o.output = argv[++i];
        } else if (a.rfind("--schedule=", 0) == 0) o.schedule = a.substr(11);
        else if (a == "--dry-run") o.dry_run = true;
        else if (a == "--support=reference") o.reference_support = true;
        else if (a == "--support=own") o.reference_support = false;
        else if (a == "--fusion=eth") o.fusion = -1;
        else if (a == "--fusion=tat_intermediate") o.fusion = DPE_TAT_INTERMEDIATE;
        else if (a == "--fusion=tat_advanced") o.fusion = DPE_TAT_ADVANCED;
        else if (!gpu_set && !a.empty() && a[0] != '-') {
            o.gpu = std::atoi(a.c_str());
            gpu_set = true;
        } else throw std::runtime_error("Unknown option: " + a);
    }
    return o;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "USAGE: dpe_driver input_folder [gpu_index] [--output=output_folder] [--vis=none|final|all] "
                     "[--checkpoint=none|final|all] [--profile=on|off] [--schedule=file] "
                     "[--support=own|reference] [--fusion=eth|tat_intermediate|tat_advanced] [--dry-run]\n";
        return EXIT_FAILURE;
    }
    Driver d;
    try {
        d.opt = parse_options(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << e.what() << "\n";
        return EXIT_FAILURE;
    }
    d.dense = argv[1];
    d.output = d.opt.output.empty() ? d.dense + "/DPE" : d.opt.output;
    if (!exists(d.dense + "/images") || !exists(d.dense + "/cams") || !exists(d.dense + "/pair.txt")) {
        std::cerr << "Invalid input folder; expected images/, cams/, and pair.txt under: " << d.dense << "\n";
        return EXIT_FAILURE;
    }
    if (d.opt.vis != Mode::NONE) std::cout << "Note: --vis images are not produced by this driver\n";
    d.cache.image_dir = d.dense + "/images";
    d.cache.cam_dir = d.dense + "/cams";
    try {
        d.load_problems();
        d.check_images();
        int min_consistent = 1;  // the reference rule (DPE.cpp:1519)
        const int rounds = d.round_num();
        std::vector<CallSpec> calls =
            d.opt.schedule.empty() ? reference_schedule(rounds) : read_schedule(d.opt.schedule, min_consistent);
        int max_scale = 1;
        for (const CallSpec &c : calls) max_scale = std::max(max_scale, c.scale);
        d.max_scale_size = d.opt.schedule.empty() ? (1 << (rounds - 1)) : max_scale;
        std::cout << "There are " << d.problems.size() << " problems needed to be processed!\n"
                  << "Round nums: " << rounds << ", calls per problem: " << calls.size() << std::endl;
        if (d.opt.dry_run) {
            for (const Problem &p : d.problems) {
                std::cout << "problem " << p.index << " ref " << p.ref << " srcs";
                for (int s : p.srcs) std::cout << " " << s;
                std::cout << "\n";
            }
            int it = 0;
            for (const CallSpec &c : calls)
                std::cout << "call " << it++ << " scale=" << c.scale << " state=" << state_name(c.state)
                          << " iters=" << c.iters << " apd=" << c.apd << " edge=" << c.edge << " geom=" << c.geom
                          << " peak=" << c.peak << " ransac=" << c.ransac << " rotate=" << c.rotate
                          << " final=" << (c.final_call ? 1 : 0) << "\n";
            std::cout << "fusion "
                      << (d.opt.fusion < 0 ? "eth" : d.opt.fusion == DPE_TAT_INTERMEDIATE ? "tat_intermediate"
                                                                                         : "tat_advanced")
                      << " min_consistent=" << min_consistent << " max_scale_size=" << d.max_scale_size
                      << std::endl;
            return EXIT_SUCCESS;
        }
        make_dirs(d.output + "/views");
        for (const Problem &p : d.problems) make_dirs(p.result_folder);
        if (dpe_create(d.opt.gpu, &d.ctx) != DPE_OK) {
            std::cerr << "dpe_create: " << dpe_last_error(nullptr) << "\n";
            return EXIT_FAILURE;
        }
        std::cout << "Input folder : " << d.dense << "\nOutput folder: " << d.output << std::endl;
        // edge precompute for every scale the schedule visits (main.cpp:373-383)
        auto te = std::chrono::steady_clock::now();
        std::vector<int> scales;
        for (const CallSpec &c : calls)
            if (std::find(scales.begin(), scales.end(), c.scale) == scales.end()) scales.push_back(c.scale);
        if (std::find(scales.begin(), scales.end(), d.max_scale_size) == scales.end()) scales.push_back(d.max_scale_size);
        for (int s : scales)
            for (const Problem &p : d.problems) d.problem_edges(p, s);
        d.prof.edge_ms += ms_since(te);
        // the schedule: each call group runs over every problem (main.cpp:390-457)
        int iteration = 0;
        for (const CallSpec &c : calls) {
            for (const Problem &p : d.problems) {
                const bool cp = d.opt.checkpoint == Mode::ALL || (d.opt.checkpoint == Mode::FINAL && c.final_call);
                d.process(p, c, iteration, cp);
            }
            iteration++;
        }
        std::cout << "[DPE Progress] patchmatch done; starting fusion" << std::endl;
        auto tf = std::chrono::steady_clock::now();
        d.fuse(min_consistent);
        d.prof.fusion_ms += ms_since(tf);
        if (d.opt.profile) {
            const Profile &p = d.prof;
            std::cout << "\n[DPE Profile Summary]\n"
                      << "  process calls       : " << p.process_calls << "\n"
                      << "  edge precompute     : " << p.edge_ms / 1000.0 << " s\n"
                      << "  input/cache lookup  : " << p.input_ms / 1000.0 << " s\n"
                      << "  support init        : " << p.support_ms / 1000.0 << " s\n"
                      << "  PatchMatch calls    : " << p.gpu_call_ms / 1000.0 << " s\n"
                      << "  GPU kernels         : " << p.gpu_kernel_ms / 1000.0 << " s\n"
                      << "  result assembly     : " << p.result_ms / 1000.0 << " s\n"
                      << "  checkpoints         : " << p.checkpoint_ms / 1000.0 << " s\n"
                      << "  fusion              : " << p.fusion_ms / 1000.0 << " s\n";
        }
    } catch (const std::exception &e) {
        std::cerr << e.what() << "\n";
        if (d.ctx) dpe_destroy(d.ctx);
        return EXIT_FAILURE;
    }
    dpe_destroy(d.ctx);
    std::cout << "All done\n";
    return EXIT_SUCCESS;
}
