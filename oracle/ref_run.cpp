// Test infrastructure only: times the compiled reference on a raw input file.
//
// Built against the reference headers (include/hpdbscan.h of
// Markus-Goetz/hpdbscan) with the reference's flags (CMakeLists.txt:12) into
// oracle/_ref/hpdbscan_ref_run.  bench.py's CPU-baseline leg runs it as a
// child process so that its OpenMP pool and its floating-point environment
// (the -Ofast build flushes denormals) stay out of the GPU process.
//
// usage: hpdbscan_ref_run <f32|f64> <input.bin> <n> <d> <eps> <minPts>
//                         <threads> <reps> [labels_out.bin]
// input.bin holds n*d values row-major; labels_out.bin receives n int64.
// Prints one JSON line: {"times_s": [...], "median_s": m, "threads": t}.
#include <unordered_set>  // hpdbscan.h:215 uses it under WITH_OUTPUT only
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hpdbscan.h"

template <typename T>
static int run(const char* path, int n, int d, float eps, size_t min_pts, int threads, int reps,
               const char* out_path) {
    std::vector<T> data(static_cast<size_t>(n) * static_cast<size_t>(d));
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        return 2;
    }
    size_t got = std::fread(data.data(), sizeof(T), data.size(), f);
    std::fclose(f);
    if (got != data.size()) {
        std::fprintf(stderr, "short read: %zu of %zu\n", got, data.size());
        return 2;
    }
    std::vector<double> times;
    Clusters labels;
    for (int r = 0; r < reps; ++r) {
        HPDBSCAN h(eps, min_pts);
        double t0 = omp_get_wtime();
        labels = h.cluster<T>(data.data(), n, d, threads);
        times.push_back(omp_get_wtime() - t0);
    }
    if (out_path) {
        FILE* o = std::fopen(out_path, "wb");
        if (!o) return 2;
        std::fwrite(labels.data(), sizeof(int64_t), labels.size(), o);
        std::fclose(o);
    }
    std::vector<double> sorted = times;
    std::sort(sorted.begin(), sorted.end());
    std::printf("{\"times_s\": [");
    for (size_t i = 0; i < times.size(); ++i) std::printf("%s%.6f", i ? ", " : "", times[i]);
    std::printf("], \"median_s\": %.6f, \"threads\": %d}\n", sorted[sorted.size() / 2], threads);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 9) {
        std::fprintf(stderr,
                     "usage: %s <f32|f64> <input.bin> <n> <d> <eps> <minPts> <threads> <reps> "
                     "[labels_out.bin]\n",
                     argv[0]);
        return 1;
    }
    const std::string dtype = argv[1];
    const int n = std::atoi(argv[3]);
    const int d = std::atoi(argv[4]);
    const float eps = std::strtof(argv[5], nullptr);
    const size_t min_pts = std::strtoull(argv[6], nullptr, 10);
    const int threads = std::atoi(argv[7]);
    const int reps = std::max(1, std::atoi(argv[8]));
    const char* out = argc > 9 ? argv[9] : nullptr;
    if (dtype == "f32") return run<float>(argv[2], n, d, eps, min_pts, threads, reps, out);
    if (dtype == "f64") return run<double>(argv[2], n, d, eps, min_pts, threads, reps, out);
    std::fprintf(stderr, "unsupported dtype %s\n", dtype.c_str());
    return 1;
}
