// ods -- command-line driver of the MI355X RTOD detector (C++ over the C-ABI).
//
// Counterpart of the reference executable optixScan: main (optixScan.cpp:
// 1128-1190), parse_args (:149-212), read_data (:95-147) and the Timer report
// (timer.h:102-117).  Same flags; the reference's compile-time DIMENSION,
// OPTIMIZATION and UPDATE_GAS_TYPE become --dim, --optimization and
// --update_gas_type.  Input: text, one point per line, "%lf" (1-D) or
// comma-separated coordinates (2-D/3-D), like read_data; or --gen for the
// seeded synthetic streams of the package's generator library (the measured
// streams: --gen walk --seed 2 --n 1048572 is the 1-D W=100k stream).  Runs
// `--warmup` untimed passes over the stream (reference: 10) and one timed
// pass, one rtod_slide call per slide or (--pipelined) all slides in one
// rtod_slides call; prints the per-slide averages.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "rtod.h"

// the package's synthetic-stream generators (rtod_amd/host/synth.c)
extern "C" {
void rtod_synth_walk(uint64_t seed, int64_t n, int dim, double start, double sigma,
                     double p_spike, double spike_lo, double spike_hi, float *out);
void rtod_synth_gmm(uint64_t seed, int64_t n, int dim, int ncomp, double lo, double hi,
                    double sig_lo, double sig_hi, double noise_frac, double drift, float *out);
}

namespace {

struct Args {
    std::string file;
    long long n = 100000000; // reference default DATA_N 1e8 (optixScan.h:4-6)
    int window = 0, slide = 0, K = 0, dim = 1;
    double R = 0.0;
    long long start_copy_pos = 0;
    int launch_ray_num = 0;
    int optimization = 2;    // the reference's OPTIMIZATION (optixScan.h:22-24)
    int update_gas_type = 1; // the reference's UPDATE_GAS_TYPE (optixScan.h:18-20)
    bool pipelined = false;
    bool pinned = false;
    int warmup = 10;
    std::string index = "auto", sched = "plain", gen, out;
    uint64_t seed = 1;
    bool phase_times = false;
};

[[noreturn]] void usage(const char *argv0)
{
    std::cerr << "Usage  : " << argv0 << " [options]\n"
              << "Options: --file | -f <filename>   input stream (one point per line)\n"
              << "         --n <int>                 number of points to read\n"
              << "         --window <int> --slide <int> --R <float> --K <int>\n"
              << "         --dim <1|2|3>             point dimension (default 1)\n"
              << "         --start_copy_pos <int>    skip this many slides of input\n"
              << "         --launch_ray_num <int>    rays cast (--optimization 0), ray i from slot i mod W\n"
              << "         --optimization 0|1|2      0 no grid filter, 1 grid filter only, 2 grid\n"
              << "                                   filter + index over the candidates (default)\n"
              << "         --update_gas_type 0|1     0 refit the whole-window LBVH, 1 rebuild (default)\n"
              << "         --index auto|runs|stencil|lbvh --sched plain|learned\n"
              << "         --pipelined               the timed pass as one pipelined call\n"
              << "         --pinned                  hold the stream in pinned memory (rtod_host_alloc):\n"
              << "                                   the kernels read the slides in place\n"
              << "         --warmup <int>            untimed passes (default 10)\n"
              << "         --gen walk|gmm --seed <u64>  synthetic stream instead of --file\n"
              << "         --out <file>              write each slide's outlier slots\n"
              << "         --phase-times             per-phase device timing\n";
    std::exit(1);
}

Args parse_args(int argc, char **argv)
{
    Args a;
    auto need = [&](int &i) -> const char * {
        if (i + 1 >= argc)
            usage(argv[0]);
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        if (s == "--help" || s == "-h")
            usage(argv[0]);
        else if (s == "--file" || s == "-f")
            a.file = need(i);
        else if (s == "--n")
            a.n = std::stoll(need(i));
        else if (s == "--window")
            a.window = std::stoi(need(i));
        else if (s == "--slide")
            a.slide = std::stoi(need(i));
        else if (s == "--R")
            a.R = std::stod(need(i));
        else if (s == "--K")
            a.K = std::stoi(need(i));
        else if (s == "--dim")
            a.dim = std::stoi(need(i));
        else if (s == "--start_copy_pos")
            a.start_copy_pos = std::stoll(need(i));
        else if (s == "--launch_ray_num")
            a.launch_ray_num = std::stoi(need(i));
        else if (s == "--optimization")
            a.optimization = std::stoi(need(i));
        else if (s == "--update_gas_type")
            a.update_gas_type = std::stoi(need(i));
        else if (s == "--pipelined")
            a.pipelined = true;
        else if (s == "--pinned")
            a.pinned = true;
        else if (s == "--index")
            a.index = need(i);
        else if (s == "--sched")
            a.sched = need(i);
        else if (s == "--warmup")
            a.warmup = std::stoi(need(i));
        else if (s == "--gen")
            a.gen = need(i);
        else if (s == "--seed")
            a.seed = std::stoull(need(i));
        else if (s == "--out")
            a.out = need(i);
        else if (s == "--phase-times")
            a.phase_times = true;
        else {
            std::cerr << "Unknown option '" << s << "'\n";
            usage(argv[0]);
        }
    }
    return a;
}

// read_data: whole file into memory (optixScan.cpp:95-147); values parsed
// as double and stored as float, the build's coordinate type.
bool read_data(const Args &a, std::vector<float> &pts)
{
    std::ifstream fin(a.file);
    if (!fin.is_open()) {
        std::cerr << "Fail to open [" << a.file << "]!" << std::endl;
        return false;
    }
    std::string line;
    pts.clear();
    long long rows = 0;
    while (rows < a.n && std::getline(fin, line)) {
        if (line.empty())
            continue;
        for (char &ch : line)
            if (ch == ',')
                ch = ' ';
        std::istringstream is(line);
        double v[3] = {0, 0, 0};
        int got = 0;
        while (got < a.dim && (is >> v[got]))
            ++got;
        if (got != a.dim) {
            std::cerr << "line " << rows + 1 << ": expected " << a.dim << " values\n";
            return false;
        }
        for (int d = 0; d < a.dim; ++d)
            pts.push_back((float)v[d]);
        ++rows;
    }
    return true;
}

// the package's seeded synthetic streams with the measured workloads'
// parameters (rtod_amd/synth.py): a random walk from 10 with N(0, 0.05)
// steps and 0.2% spikes of +-U(2, 5); a mixture of 8 Gaussians in [0, 100]^d,
// sigma U(0.5, 3), 0.5% uniform noise, centres drifting by 5
void generate(const Args &a, std::vector<float> &pts)
{
    pts.resize((size_t)a.n * a.dim);
    if (a.gen == "walk")
        rtod_synth_walk(a.seed, a.n, a.dim, 10.0, 0.05, 0.002, 2.0, 5.0, pts.data());
    else
        rtod_synth_gmm(a.seed, a.n, a.dim, 8, 0.0, 100.0, 0.5, 3.0, 0.005, 5.0, pts.data());
}

long vm_rss_kb()
{
    FILE *f = std::fopen("/proc/self/status", "r");
    if (!f)
        return -1;
    char line[256];
    long kb = -1;
    while (std::fgets(line, sizeof line, f))
        if (std::strncmp(line, "VmRSS:", 6) == 0) {
            kb = std::atol(line + 6);
            break;
        }
    std::fclose(f);
    return kb;
}

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        rtod_status s_ = (x);                                                                     \
        if (s_ != RTOD_OK) {                                                                      \
            std::cerr << #x << ": " << rtod_status_str(s_) << ": " << rtod_last_error() << "\n";  \
            return 2;                                                                             \
        }                                                                                         \
    } while (0)

} // namespace

int main(int argc, char **argv)
{
    Args a = parse_args(argc, argv);
    if (a.window <= 0 || a.slide <= 0 || a.K <= 0 || !(a.R > 0)) {
        std::cerr << "--window, --slide, --R and --K are required\n";
        usage(argv[0]);
    }
    std::vector<float> pts;
    if (!a.gen.empty()) {
        if (a.n >= 100000000)
            a.n = 1048572;
        generate(a, pts);
    } else if (!read_data(a, pts)) {
        return 1;
    }
    const long long rows = (long long)(pts.size() / a.dim);
    const long long skip = a.start_copy_pos * a.slide; // optixScan.cpp:1136
    if (rows - skip < a.window) {
        std::cerr << "not enough points for one window\n";
        return 1;
    }
    const float *base = pts.data() + (size_t)skip * a.dim;
    // --pinned: the stream copied once into a buffer from rtod_host_alloc,
    // whose slides rtod_slide / rtod_slides read in place (no staging copy)
    float *pin = nullptr;
    if (a.pinned) {
        CHECK(rtod_host_alloc((int64_t)(pts.size() * sizeof(float)), (void **)&pin));
        std::memcpy(pin, pts.data(), pts.size() * sizeof(float));
        base = pin + (size_t)skip * a.dim;
    }
    const long long avail = rows - skip;
    const long long T = std::min<long long>(10000, (avail - a.window) / a.slide);

    rtod_config cfg{};
    cfg.dim = a.dim;
    cfg.window = a.window;
    cfg.slide = a.slide;
    cfg.R = (float)a.R;
    cfg.K = a.K;
    cfg.index_kind = a.index == "lbvh"      ? RTOD_INDEX_LBVH
                     : a.index == "stencil" ? RTOD_INDEX_CELL_STENCIL
                     : a.index == "runs"    ? RTOD_INDEX_SLIDE_RUNS
                                            : RTOD_INDEX_AUTO;
    cfg.sched_kind = a.sched == "learned" ? RTOD_SCHED_LEARNED : RTOD_SCHED_PLAIN;
    cfg.miss_threshold = 0.25f;
    cfg.flags = a.phase_times ? RTOD_FLAG_PHASE_TIMES : 0;
    cfg.device = 0;
    cfg.ablation = a.optimization == 0   ? RTOD_ABL_NO_FILTER
                   : a.optimization == 1 ? RTOD_ABL_FILTER_ONLY
                                         : RTOD_ABL_NONE;
    cfg.refit = a.update_gas_type == 0 ? 1 : 0;
    cfg.launch_ray_num = a.launch_ray_num;

    std::cout << "Dimension: " << a.dim << "\nR: " << a.R << ", R2: " << (float)a.R * (float)a.R
              << "\nK: " << a.K << "\nWindow size: " << a.window << "\nSlide size: " << a.slide
              << "\nData num: " << rows << "\nIndex: " << a.index << "\nSchedule: " << a.sched
              << "\nOptimization: " << a.optimization << "\nUpdate GAS type: " << a.update_gas_type
              << "\n";
    if (a.launch_ray_num)
        std::cout << "Launch ray num: " << a.launch_ray_num << "\n";

    const long rss0 = vm_rss_kb();
    rtod_ctx *ctx = nullptr;
    CHECK(rtod_create(&cfg, &ctx));
    const size_t cap = (size_t)std::max(a.window, a.launch_ray_num);
    std::vector<int32_t> outl(cap);
    FILE *fo = a.out.empty() ? nullptr : std::fopen(a.out.c_str(), "w");
    rtod_times acc{};
    int64_t total_out = 0;
    double pipe_ms = 0.0;
    auto write_slide = [&](long long t, int n, const int32_t *l) {
        if (!fo)
            return;
        std::fprintf(fo, "%lld %d", t, n);
        for (int i = 0; i < n; ++i)
            std::fprintf(fo, " %d", l[i]);
        std::fprintf(fo, "\n");
    };
    for (int pass = 0; pass <= a.warmup; ++pass) {
        const bool timed = pass == a.warmup;
        CHECK(rtod_prime(ctx, base, a.window));
        acc = rtod_times{};
        total_out = 0;
        if (a.pipelined && T > 0) { // all slides in one call, from host memory
            std::vector<int32_t> offs((size_t)T + 1), slots((size_t)T * cap);
            const auto t0 = std::chrono::steady_clock::now();
            CHECK(rtod_slides(ctx, base + (size_t)a.window * a.dim, (int32_t)T, offs.data(),
                              slots.data(), (int64_t)slots.size()));
            const auto t1 = std::chrono::steady_clock::now();
            if (!timed)
                continue;
            pipe_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            for (long long t = 1; t <= T; ++t) {
                const int n = offs[t] - offs[t - 1];
                total_out += n;
                write_slide(t, n, slots.data() + offs[t - 1]);
            }
            continue;
        }
        for (long long t = 1; t <= T; ++t) {
            rtod_times tm{};
            int32_t n = 0;
            CHECK(rtod_slide(ctx, base + (size_t)(a.window + (t - 1) * a.slide) * a.dim,
                             outl.data(), &n, &tm));
            if (!timed)
                continue;
            acc.copy_h2d += tm.copy_h2d;
            acc.prepare_cell += tm.prepare_cell;
            acc.build_index += tm.build_index;
            acc.detect_outlier += tm.detect_outlier;
            acc.copy_d2h += tm.copy_d2h;
            acc.total += tm.total;
            acc.launch += tm.launch;
            total_out += n;
            write_slide(t, n, outl.data());
        }
    }
    if (fo)
        std::fclose(fo);
    const double Td = T > 0 ? (double)T : 1.0;
    rtod_stats st{};
    CHECK(rtod_get_stats(ctx, &st));
    if (a.pipelined) {
        std::cout << "\n###########   Time  ##########\n"
                  << "[Time] pipelined slides from host memory" << (a.pinned ? " (pinned)" : "")
                  << ": " << pipe_ms / Td
                  << " ms per slide\n"
                  << "##############################\n\n";
    } else {
        std::cout << "\n###########   Time  ##########\n"
                  << "[Time] copy new points h2d: " << acc.copy_h2d / Td << " ms\n"
                  << "[Time] copy filtered points h2d: 0 ms\n"
                  << "[Time] copy outlier d2h: " << acc.copy_d2h / Td << " ms\n"
                  << "[Time] prepare cell: " << acc.prepare_cell / Td << " ms\n"
                  << "[Time] build BVH: " << acc.build_index / Td << " ms\n"
                  << "[Time] detect outlier: " << acc.detect_outlier / Td << " ms\n"
                  << "[Time] launch and completion: " << acc.launch / Td << " ms\n"
                  << "[Time] total time for a slide: " << acc.total / Td << " ms\n"
                  << "##############################\n\n";
    }
    std::cout << "[Mem] Host cpu memory used(MB): " << (vm_rss_kb() - rss0) / 1024.0 << "\n"
              << "[Mem] Device memory used for data(MB): " << st.device_bytes / 1048576.0 << "\n"
              << "[Mem] Pinned host memory(MB): " << st.host_pinned_bytes / 1048576.0 << "\n";
    std::cout << "BVH Node / slide: " << (st.slides ? st.candidates_total / st.slides : 0) << "\n"
              << "Outliers / slide: " << total_out / Td << "\n"
              << "Pair tests / slide: " << (st.slides ? st.pair_tests_total / st.slides : 0)
              << "\n"
              << "Hit / slide: " << (st.slides ? st.hits_total / st.slides : 0) << "\n";
    if (cfg.sched_kind == RTOD_SCHED_LEARNED)
        std::cout << "Predictor misses: " << st.predictor_misses << " / " << st.predictor_points
                  << ", fallbacks: " << st.fallbacks << "\n";
    rtod_destroy(ctx);
    CHECK(rtod_host_free(pin));
    return 0;
}
