// nanogi-compatible command line over the C ABI.
//
// Mirrors Run() / main() of src/nanogi.cpp:1993-2151: the same option names,
// defaults and positional order (`nanogi [options] <renderer> <scene>
// <result> <width> <height>`), usage on --help or no arguments (exit 0),
// scene loaded with aspect W/H (:2069), render, SaveImage (:2097-2114).
// Renderers other than pt/ptdirect/lt/ltdirect/bdpt fail with "unsupported renderer".
// Additional options (not in the reference): --spp, --seed, --rr-depth,
// --rr-prob, --precision, --engine, --paths-in-flight, --autotune.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ngi.h"

namespace {

struct Opt {
    const char* name;
    char shortname;
    const char* def;   // nullptr: no default
    const char* help;
    bool flag;
};

const Opt kOpts[] = {
    {"help", 0, nullptr, "Display help message", true},
    {"scene", 'i', nullptr, "Scene file", false},
    {"result", 'o', "render.hdr", "Rendered result", false},
    {"renderer", 'r', nullptr, "Rendering technique", false},
    {"num-samples", 'n', "10000000", "Number of samples", false},
    {"max-num-vertices", 'm', "-1", "Maximum number of vertices", false},
    {"width", 'w', "1280", "Width of the rendered image", false},
    {"height", 'h', "720", "Height of the rendered image", false},
    {"num-threads", 'j', nullptr, "Number of threads (host side only)", false},
    {"grain-size", 0, "10000", "Grain size (paths per time-budget chunk / 1000)", false},
    {"progress-update-interval", 0, "100000", "Progress update interval", false},
    {"render-time", 't', "-1", "Render time in seconds (-1 to use # of samples)", false},
    {"progress-image-update-interval", 0, "-1", "Progress image update interval (-1: disable)", false},
    {"progress-image-update-format", 0, "progress/{{count}}.png", "Progress image update format string\n - {{count}}: image count", false},
    // Additions.
    {"spp", 0, nullptr, "Samples per pixel (per-pixel stratified sampling instead of -n)", false},
    {"seed", 0, "1", "Random seed", false},
    {"rr-depth", 0, "1", "Russian roulette from this vertex count (-1: off)", false},
    {"rr-prob", 0, "0.5", "Russian roulette continuation probability", false},
    {"precision", 0, "fp32", "fp32 (FP64 film) | fp32all | fp64", false},
    {"engine", 0, "persistent", "persistent | wavefront | megakernel", false},
    {"paths-in-flight", 0, "0", "Concurrent paths: wavefront queue capacity or persistent lanes (0: auto)", false},
    {"autotune", 0, nullptr, "Time the engines on a sample of this many paths and render with the fastest", false},
};
const char* kPositional[] = {"renderer", "scene", "result", "width", "height"};

void usage() {
    printf("Usage: nanogi [options] <renderer> <scene> <result> <width> <height>\nAllowed options:\n");
    for (const Opt& o : kOpts) {
        char left[96];
        if (o.shortname) snprintf(left, sizeof(left), "  -%c [ --%s ]%s", o.shortname, o.name, o.flag ? "" : " arg");
        else snprintf(left, sizeof(left), "  --%s%s", o.name, o.flag ? "" : " arg");
        printf("%-44s %s%s%s%s\n", left, o.help, o.def ? " (=" : "", o.def ? o.def : "", o.def ? ")" : "");
    }
}

const Opt* find_long(const std::string& n) {
    for (const Opt& o : kOpts)
        if (n == o.name) return &o;
    return nullptr;
}
const Opt* find_short(char c) {
    for (const Opt& o : kOpts)
        if (o.shortname == c) return &o;
    return nullptr;
}

bool parse(int argc, char** argv, std::map<std::string, std::string>& vm, std::string& err) {
    std::vector<std::string> positional;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        const Opt* o = nullptr;
        std::string val;
        bool has_val = false;
        if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            std::string n = a.substr(2);
            const size_t eq = n.find('=');
            if (eq != std::string::npos) {
                val = n.substr(eq + 1);
                n = n.substr(0, eq);
                has_val = true;
            }
            o = find_long(n);
            if (!o) {
                err = "unrecognised option '" + a + "'";
                return false;
            }
        } else if (a.size() >= 2 && a[0] == '-' && !(a[1] >= '0' && a[1] <= '9')) {
            o = find_short(a[1]);
            if (!o) {
                err = "unrecognised option '" + a + "'";
                return false;
            }
            if (a.size() > 2) val = a.substr(2), has_val = true;
        } else {
            positional.push_back(a);
            continue;
        }
        if (o->flag) {
            vm[o->name] = "1";
            continue;
        }
        if (!has_val) {
            if (i + 1 >= argc) {
                err = std::string("the required argument for option '--") + o->name + "' is missing";
                return false;
            }
            val = argv[++i];
        }
        vm[o->name] = val;
    }
    // Positional values fill renderer, scene, result, width, height in order,
    // skipping the ones already given as named options.
    size_t slot = 0;
    for (const std::string& v : positional) {
        while (slot < 5 && vm.count(kPositional[slot])) ++slot;
        if (slot >= 5) {
            err = "too many positional options";
            return false;
        }
        vm[kPositional[slot++]] = v;
    }
    for (const Opt& o : kOpts)
        if (o.def && !vm.count(o.name)) vm[o.name] = o.def;
    return true;
}

std::string expand_count(const std::string& fmt, long long count) {
    char c[32];
    snprintf(c, sizeof(c), "%010lld", count);
    std::string s = fmt;
    const size_t p = s.find("{{count}}");
    if (p != std::string::npos) s.replace(p, 9, c);
    return s;
}

int run(int argc, char** argv) {
    std::map<std::string, std::string> vm;
    std::string err;
    if (!parse(argc, argv, vm, err)) {
        fprintf(stderr, "ERROR : %s\n", err.c_str());
        return EXIT_FAILURE;
    }
    if (vm.count("help") || argc == 1) {
        usage();
        return EXIT_SUCCESS;   // Run() returns 1 (true) here (src/nanogi.cpp:2030-2034)
    }
    if (!vm.count("renderer")) {
        fprintf(stderr, "ERROR : the option '--renderer' is required but missing\n");
        return EXIT_FAILURE;
    }
    if (!vm.count("scene")) {
        fprintf(stderr, "ERROR : the option '--scene' is required but missing\n");
        return EXIT_FAILURE;
    }
    ngi_render_params p;
    ngi_render_params_default(&p);
    const int r = ngi_renderer_from_string(vm["renderer"].c_str());
    if (r < NGI_RENDERER_PT || r > NGI_RENDERER_BDPT) {
        fprintf(stderr, "ERROR : unsupported renderer '%s' (this build provides pt, ptdirect, lt, ltdirect and bdpt)\n", vm["renderer"].c_str());
        return EXIT_FAILURE;
    }
    p.renderer = r;
    p.width = atoi(vm["width"].c_str());
    p.height = atoi(vm["height"].c_str());
    p.num_samples = atoll(vm["num-samples"].c_str());
    p.max_num_vertices = atoi(vm["max-num-vertices"].c_str());
    p.seed = (uint32_t)strtoul(vm["seed"].c_str(), nullptr, 10);
    p.rr_depth = atoi(vm["rr-depth"].c_str());
    p.rr_prob = atof(vm["rr-prob"].c_str());
    p.paths_in_flight = atoi(vm["paths-in-flight"].c_str());
    const std::string prec = vm["precision"], eng = vm["engine"];
    p.precision = prec == "fp64" ? NGI_PRECISION_FP64_ALL : (prec == "fp32all" ? NGI_PRECISION_FP32_ALL : NGI_PRECISION_FP32_FP64ACC);
    p.engine = eng == "wavefront" ? NGI_ENGINE_WAVEFRONT : (eng == "megakernel" ? NGI_ENGINE_MEGAKERNEL : NGI_ENGINE_PERSISTENT);
    if (p.precision == NGI_PRECISION_FP64_ALL && p.engine != NGI_ENGINE_MEGAKERNEL && p.engine != NGI_ENGINE_PERSISTENT)
        p.engine = NGI_ENGINE_PERSISTENT;
    int spp = 0;
    if (vm.count("spp")) {
        spp = atoi(vm["spp"].c_str());
        p.sampling = NGI_SAMPLING_PER_PIXEL;
        p.spp = spp;
    }
    if (p.width <= 0 || p.height <= 0) {
        fprintf(stderr, "ERROR : invalid image size\n");
        return EXIT_FAILURE;
    }
    printf("nanogi (MI355X pt/ptdirect/lt/ltdirect/bdpt)\n");
    ngi_scene* scene = nullptr;
    printf("Loading scene\n");
    if (ngi_scene_load_yaml(vm["scene"].c_str(), (double)p.width / p.height, &scene) != NGI_OK) {
        fprintf(stderr, "ERROR : %s\n", ngi_last_error());
        return EXIT_FAILURE;
    }
    if (vm.count("autotune")) {
        ngi_tune_trial trials[16];
        int n = 0;
        if (ngi_autotune(scene, &p, atoll(vm["autotune"].c_str()), trials, 16, &n) != NGI_OK) {
            fprintf(stderr, "ERROR : %s\n", ngi_last_error());
            ngi_scene_destroy(scene);
            return EXIT_FAILURE;
        }
        const char* names[] = {"megakernel", "wavefront", "persistent"};
        for (int i = 0; i < std::min(n, 16); ++i)
            printf("Autotune: %-10s paths in flight %8d: %.3f ms\n", names[trials[i].engine], trials[i].paths_in_flight, trials[i].ms);
        printf("Autotune: using %s, paths in flight %d\n", names[p.engine], p.paths_in_flight);
    }
    printf("Rendering\n");
    const double render_time = atof(vm["render-time"].c_str());
    const double img_interval = atof(vm["progress-image-update-interval"].c_str());
    const std::string img_fmt = vm["progress-image-update-format"];
    const long long grain = atoll(vm["grain-size"].c_str());
    std::vector<double> film((size_t)p.width * p.height * 3);
    ngi_session* sess = nullptr;
    int rc = EXIT_SUCCESS;
    // Time-budget mode renders chunks of GrainSize*1000 paths (src/nanogi.cpp:243).
    if (render_time > 0 && !spp) p.num_samples = std::max(1LL, grain * 1000);
    if (ngi_session_create(scene, &p, &sess) != NGI_OK) {
        fprintf(stderr, "ERROR : %s\n", ngi_last_error());
        ngi_scene_destroy(scene);
        return EXIT_FAILURE;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto last_img = t0;
    long long img_count = 0;
    auto elapsed = [&](std::chrono::steady_clock::time_point a) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };
    for (;;) {
        if (ngi_session_render(sess, spp ? spp : 1) != NGI_OK || ngi_session_sync(sess) != NGI_OK) {
            fprintf(stderr, "ERROR : %s\n", ngi_last_error());
            rc = EXIT_FAILURE;
            break;
        }
        if (img_interval > 0 && elapsed(last_img) > img_interval) {
            ngi_session_get_film(sess, film.data(), nullptr);
            const std::string path = expand_count(img_fmt, ++img_count);
            printf("Saving progress: %s\n", path.c_str());
            if (ngi_image_save(path.c_str(), film.data(), p.width, p.height) != NGI_OK) fprintf(stderr, "ERROR : %s\n", ngi_last_error());
            last_img = std::chrono::steady_clock::now();
        }
        if (render_time <= 0 || elapsed(t0) > render_time) break;
    }
    ngi_stats st;
    memset(&st, 0, sizeof(st));
    if (rc == EXIT_SUCCESS) {
        ngi_session_get_stats(sess, &st);
        printf("Progress: 100.0%%\n# of samples: %lld\n", (long long)st.samples);
        printf("Elapsed time: %.3f\n", elapsed(t0));
        if (ngi_session_get_film(sess, film.data(), nullptr) != NGI_OK) rc = EXIT_FAILURE;
    }
    ngi_session_destroy(sess);
    ngi_scene_destroy(scene);
    if (rc != EXIT_SUCCESS) return rc;
    printf("Saving rendered image\n");
    if (ngi_image_save(vm["result"].c_str(), film.data(), p.width, p.height) != NGI_OK) {
        fprintf(stderr, "ERROR : %s\n", ngi_last_error());
        return EXIT_FAILURE;
    }
    printf("Successfully saved to %s\n", vm["result"].c_str());
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }
