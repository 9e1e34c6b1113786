// Command-line tool: clusters a 2-D HDF5 dataset on the GPU and writes the
// labels back as an HDF5 dataset.  Options, defaults and error behaviour
// follow the reference's src/cli.cpp:42-137:
//
//   -m, --minPoints N        density threshold            (default 2)
//   -e, --epsilon X          spatial search radius        (default 0.1)
//   -t, --threads N          accepted; the GPU path uses no host threads
//                                                         (default: hardware threads)
//   -i, --input PATH         input file                   (default data.h5)
//   -o, --output PATH        output file                  (default data.h5)
//       --input-dataset S    input dataset name           (default DATA)
//       --output-dataset S   output dataset name          (default CLUSTERS)
//   -h, --help               this help message
//
// plus one option of its own:
//       --summary            print the reference's summary block
//                            (hpdbscan.h:213-284, shown there when built
//                            WITH_OUTPUT) and the total time
//
// Values go as "-e 0.3", "-e0.3", "--epsilon 0.3" or "--epsilon=0.3".  A
// command line that does not parse prints "Could not parse command line:"
// and the reason, worded as the reference's option parser (cxxopts 2.2)
// words it; a failed run prints the exception's message; both exit with
// status 0, as the reference's tool does (cli.cpp:55-67, :86-132).
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hpdbscan.h"

namespace {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Option {
    char short_name;
    std::string long_name;
    bool takes_value;
    std::string value;
    std::string help;
};

class Options {
public:
    Options() {
        const std::string threads = std::to_string(std::max(1u, std::thread::hardware_concurrency()));
        add('h', "help", false, "", "this help message");
        add('m', "minPoints", true, "2", "density threshold");
        add('e', "epsilon", true, "0.1", "spatial search radius");
        add('t', "threads", true, threads, "utilized threads");
        add('i', "input", true, "data.h5", "input file");
        add('o', "output", true, "data.h5", "output file");
        add(0, "input-dataset", true, "DATA", "input dataset name");
        add(0, "output-dataset", true, "CLUSTERS", "output dataset name");
        add(0, "summary", false, "", "print cluster, noise and core point counts");
    }

    void parse(int argc, char** argv) {
        for (int a = 1; a < argc; ++a) {
            const std::string arg = argv[a];
            if (arg.size() > 2 && arg[0] == '-' && arg[1] == '-') {
                const size_t eq = arg.find('=');
                const std::string name = arg.substr(2, eq == std::string::npos ? eq : eq - 2);
                Option& o = by_long(name);
                if (!o.takes_value) {
                    if (eq != std::string::npos)
                        throw ParseError("Option \u2018" + name + "\u2019 does not take an argument");
                    set_flag(o);
                } else if (eq != std::string::npos) {
                    set_value(o, arg.substr(eq + 1));
                } else {
                    if (a + 1 >= argc) throw ParseError("Option \u2018" + name + "\u2019 is missing an argument");
                    set_value(o, argv[++a]);
                }
            } else if (arg.size() > 1 && arg[0] == '-') {
                // a group of short flags; an option that takes a value takes
                // the rest of the group or the next argument
                for (size_t k = 1; k < arg.size(); ++k) {
                    Option& o = by_short(arg[k]);
                    if (!o.takes_value) {
                        set_flag(o);
                        continue;
                    }
                    if (k + 1 < arg.size()) {
                        set_value(o, arg.substr(k + 1));
                    } else {
                        if (a + 1 >= argc)
                            throw ParseError("Option \u2018" + std::string(1, arg[k]) +
                                             "\u2019 is missing an argument");
                        set_value(o, argv[++a]);
                    }
                    break;
                }
            }
            // other arguments are ignored, as by the reference's parser
        }
    }

    bool given(const std::string& name) const { return given_.count(name) > 0; }
    const std::string& value(const std::string& name) const { return find(name).value; }

    std::string help() const {
        std::string h = "Highly parallel DBSCAN clustering algorithm (MI355X)\nUsage:\n  hpdbscan [OPTION...]\n\n";
        for (const Option& o : options_) {
            std::string left = "  ";
            left += o.short_name ? std::string("-") + o.short_name + ", " : "    ";
            left += "--" + o.long_name + (o.takes_value ? " arg" : "");
            left.resize(std::max<size_t>(left.size() + 2, 28), ' ');
            h += left + o.help;
            if (o.takes_value) h += " (default: " + o.value + ")";
            h += "\n";
        }
        return h;
    }

private:
    void add(char s, const std::string& l, bool v, const std::string& def, const std::string& help) {
        options_.push_back(Option{s, l, v, def, help});
    }
    Option& by_long(const std::string& name) {
        for (Option& o : options_)
            if (o.long_name == name) return o;
        throw ParseError("Option \u2018" + name + "\u2019 does not exist");
    }
    Option& by_short(char c) {
        for (Option& o : options_)
            if (c && o.short_name == c) return o;
        throw ParseError("Option \u2018" + std::string(1, c) + "\u2019 does not exist");
    }
    const Option& find(const std::string& name) const {
        for (const Option& o : options_)
            if (o.long_name == name) return o;
        throw std::logic_error("unknown option " + name);
    }
    void set_flag(Option& o) { given_.insert(o.long_name); }
    void set_value(Option& o, const std::string& v) {
        o.value = v;
        given_.insert(o.long_name);
    }

    std::vector<Option> options_;
    std::set<std::string> given_;
};

// Whole-string conversions; a value that does not convert is a parse error
// (cxxopts' argument_incorrect_type in the reference).
size_t to_size(const std::string& s) {
    if (s.empty() || s[0] == '-' || s[0] == '+') throw ParseError("Argument \u2018" + s + "\u2019 failed to parse");
    errno = 0;
    char* end = nullptr;
    const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
    if (errno || *end) throw ParseError("Argument \u2018" + s + "\u2019 failed to parse");
    return static_cast<size_t>(v);
}

int to_int(const std::string& s) {
    errno = 0;
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || errno || *end || v < INT32_MIN || v > INT32_MAX)
        throw ParseError("Argument \u2018" + s + "\u2019 failed to parse");
    return static_cast<int>(v);
}

float to_float(const std::string& s) {
    errno = 0;
    char* end = nullptr;
    const float v = std::strtof(s.c_str(), &end);
    if (s.empty() || *end) throw ParseError("Argument \u2018" + s + "\u2019 failed to parse");
    return v;
}

// The reference's summary block (hpdbscan.h:213-284): clusters are the
// distinct |label| values, less the noise label 0 when there is noise.
void summarize(const Clusters& clusters) {
    std::set<Cluster> unique;
    size_t cluster_points = 0, noise_points = 0, core_points = 0;
    for (const Cluster c : clusters) {
        unique.insert(c < 0 ? -c : c);
        if (c == 0)
            ++noise_points;
        else
            ++cluster_points;
        if (c < 0) ++core_points;
    }
    std::cout << "Summary..." << std::endl;
    std::cout << "\tClusters:       " << (noise_points ? unique.size() - 1 : unique.size()) << std::endl
              << "\tCluster points: " << cluster_points << std::endl
              << "\tNoise points:   " << noise_points << std::endl
              << "\tCore points:    " << core_points << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
    Options opts;
    size_t min_points = 0;
    float epsilon = 0.0f;
    int threads = 0;
    try {
        opts.parse(argc, argv);
        min_points = to_size(opts.value("minPoints"));
        epsilon = to_float(opts.value("epsilon"));
        threads = to_int(opts.value("threads"));
    } catch (const ParseError& e) {
        std::cerr << "Could not parse command line:" << std::endl;
        std::cerr << e.what() << std::endl;
        return 0;
    }
    if (opts.given("help")) {
        std::cout << opts.help() << std::endl;
        return 0;
    }

    const auto t0 = std::chrono::steady_clock::now();
    Clusters clusters;
    try {
        HPDBSCAN hpdbscan(epsilon, min_points);
        clusters = hpdbscan.cluster(opts.value("input"), opts.value("input-dataset"), threads);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 0;
    }
    if (opts.given("summary")) {
        summarize(clusters);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Total time: " << sec << std::endl;
    }
    try {
        hpdbscan_amd::write_hdf5(opts.value("output"), opts.value("output-dataset"), clusters);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 0;
    }
    return 0;
}
