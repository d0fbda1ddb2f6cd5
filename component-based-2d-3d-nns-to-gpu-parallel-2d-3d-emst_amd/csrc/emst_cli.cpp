// emst_cli.cpp -- command line front end with the reference's argument
// order and outputs (emst_coalition/src/main.cpp:18-92,
// CalculateurEMST.cu:52-262, SolutionEMSTRW.cu:24-189):
//
//   emst_cli [data [solution [stats [config]]]]
//   defaults: input.svg output.svg output.stats config.cfg
//
// The config file is the reference's sectioned "key = value" format; read
// here: [coalition_global_param] functionModeChoice (4 RUN: 2D points,
// 5 RUN_3D: 3D points, 6 RUN_2DPOINT5: 3D points with cones over the xy
// plane, CalculateurEMST.h:40; every other mode is reported unsupported, as
// the reference does for modes it does not build), [global_param] traceActive
// (one stats row per round) and levelRadius.  Without a config file the mode
// follows the data suffix (.3d -> 3D, otherwise 2D) instead of the
// reference's default mode 2, which runs no EMST.
//
// Outputs: the stats file (18 columns, header when the file is empty, a row
// per round when tracing and a final row; SolutionEMSTRW.cu:85-189) and the
// solution maps (emst_write_solution).
#include <chrono>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../include/emst.h"

namespace {

std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// "[section]" headers and "key = value" lines; '#' and ';' start comments
std::map<std::string, std::string> read_config(const std::string& path, bool* found)
{
    std::map<std::string, std::string> kv;
    std::ifstream fi(path);
    *found = (bool)fi;
    std::string line, section;
    while (std::getline(fi, line)) {
        const size_t c = line.find_first_of("#;");
        if (c != std::string::npos) line = line.substr(0, c);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[' && line.back() == ']') {
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        kv[section + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return kv;
}

int get_int(const std::map<std::string, std::string>& kv, const std::string& key, int def)
{
    const auto it = kv.find(key);
    if (it == kv.end()) return def;
    try {
        return std::stoi(it->second);
    } catch (...) {
        return def;
    }
}

struct Stats {
    std::string path, bench;
    std::time_t t0 = 0;
    std::clock_t x0 = 0;
    std::chrono::steady_clock::time_point g0;
    double cum[5] = {0, 0, 0, 0, 0};     // find next closest, find min pair, connect union, flatten, terminate

    // writeHeaderStatistics: the header goes to an empty file only
    void header(std::ostream& o, bool file_is_empty)
    {
        if (!file_is_empty) return;
        o << "iteration" << std::setw(30) << "global_objectif" << std::setw(30) << "duree(s)" << std::setw(30)
          << "duree(s.xx)" << std::setw(30) << "cuda_duree(ms)" << std::setw(30) << "benchMark" << std::setw(30)
          << "size" << std::setw(30) << "length" << std::setw(30) << "timeFlatten" << std::setw(30)
          << "timeTestTermination" << std::setw(30) << "timeFindNextClosest" << std::setw(30) << "timeFindMinPair"
          << std::setw(30) << "timeConnectGraphUnion" << std::setw(30) << "timeCumulativeFindNextClosest"
          << std::setw(30) << "timeCumulativeFindMinPair" << std::setw(30) << "timeCumulativeConnectUnion"
          << std::setw(30) << "timeCumulativeFlatten" << std::setw(30) << "timeCumulativeTerminate" << std::endl;
    }

    // writeStatistics: one row, then two blank lines
    void row(std::ostream& o, int iteration, const emst_round_trace& r)
    {
        std::time_t tf;
        const double gpu_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
        o << iteration << std::setw(30) << r.size << std::setw(30) << (long long)(std::time(&tf) - t0)
          << std::setw(30) << (double)(std::clock() - x0) / CLOCKS_PER_SEC << std::setw(30) << gpu_ms
          << std::setw(30) << bench << std::setw(30) << r.size << std::setw(30) << r.length << std::setw(30)
          << r.time_flatten << std::setw(30) << r.time_termination << std::setw(30) << r.time_find_next_closest
          << std::setw(30) << r.time_find_min_pair << std::setw(30) << r.time_connect_union << std::setw(30)
          << cum[0] << std::setw(30) << cum[1] << std::setw(30) << cum[2] << std::setw(30) << cum[3]
          << std::setw(30) << cum[4] << std::endl;
        o << std::endl << std::endl;
    }

    void append(int iteration, const emst_round_trace& r)
    {
        std::ofstream fo(path, std::ios::app);
        if (!fo) {
            std::cerr << "Unable to open file " << path << " CRITICAL ERROR" << std::endl;
            return;
        }
        row(fo, iteration, r);
    }
};

bool file_empty(const std::string& path)
{
    std::ifstream fin(path);
    std::string s;
    fin >> s;
    return s.empty();
}

int fail(emst_ctx* c, const char* what, int rc)
{
    std::cerr << what << " failed (" << rc << "): " << emst_last_error(c) << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char* argv[])
{
    std::string fileData = "input.svg", fileSolution = "output.svg", fileStats = "output.stats",
                fileConfig = "config.cfg";
    if (argc >= 2) fileData = argv[1];
    if (argc >= 3) fileSolution = argv[2];
    if (argc >= 4) fileStats = argv[3];
    if (argc >= 5) fileConfig = argv[4];
    std::cout << argv[0] << " " << fileData << " " << fileSolution << " " << fileStats << " " << fileConfig
              << std::endl;

    bool have_cfg = false;
    const auto kv = read_config(fileConfig, &have_cfg);
    const size_t dot = fileData.rfind('.');
    const bool is3d = dot != std::string::npos && fileData.substr(dot) == ".3d";
    int mode = get_int(kv, "coalition_global_param.functionModeChoice", -1);
    if (mode < 0) {
        mode = is3d ? 5 : 4;
        if (!have_cfg) std::cout << "no config file: mode " << mode << " from the data suffix" << std::endl;
    }
    if (mode != 4 && mode != 5 && mode != 6) {
        std::cout << "UNSUPPORTED FUNCTIONMODE=" << mode << " !! " << std::endl;
        return 0;
    }
    const int dim = mode == 4 ? 2 : 3;
    const bool trace = get_int(kv, "global_param.traceActive", 1) != 0;

    int64_t n = 0;
    int rc = emst_read_points(fileData.c_str(), dim, nullptr, &n);
    if (rc) {
        std::cout << "erreur ouverture file: " << fileData << std::endl;
        return 1;
    }
    std::vector<float> xyz((size_t)(n * dim));
    if (n > 0 && (rc = emst_read_points(fileData.c_str(), dim, xyz.data(), &n))) {
        std::cout << "erreur lecture file: " << fileData << std::endl;
        return 1;
    }
    std::cout << "BEGIN READ " << fileData << " " << n << " points" << std::endl;

    emst_options o;
    emst_default_options(&o);
    o.dim = dim;
    o.cone_dim = mode == 6 ? 2 : 0;
    o.trace = 1;            // per-round times and lengths for the stats rows
    o.level_radius = get_int(kv, "global_param.levelRadius", 1);
    emst_ctx* c = emst_create(&o);
    if (!c) {
        std::cerr << "emst_create failed: " << emst_last_error(nullptr) << std::endl;
        return 1;
    }

    Stats st;
    st.path = fileStats;
    // benchMark column: the data file's name (the reference is run on bare
    // names; a directory part would overflow the 30-wide column)
    st.bench = fileData.substr(fileData.find_last_of('/') == std::string::npos ? 0 : fileData.find_last_of('/') + 1);
    std::time(&st.t0);
    st.x0 = std::clock();
    st.g0 = std::chrono::steady_clock::now();
    {
        // initStatisticsFile (SolutionEMSTRW.cu:85-97)
        const bool empty = file_empty(fileStats);
        std::ofstream fo(fileStats, std::ios::app);
        if (!fo) {
            std::cerr << "Unable to open file " << fileStats << " CRITICAL ERROR" << std::endl;
            emst_destroy(c);
            return 1;
        }
        st.header(fo, empty);
    }

    if ((rc = emst_set_points(c, xyz.data(), n, dim, 0))) return fail(c, "set_points", rc);
    // run(): one stats row per round while tracing (SolutionEMSTOperators.cu:146-153)
    int32_t finished = 0;
    int iteration = 0;
    emst_round_trace last{};
    float build[4] = {0, 0, 0, 0};
    while (true) {
        if ((rc = emst_step(c, &finished))) {
            const int r = fail(c, "step", rc);
            emst_destroy(c);
            return r;
        }
        if (finished) break;
        int32_t k = 0;
        emst_get_trace(c, nullptr, &k);
        std::vector<emst_round_trace> rows((size_t)k);
        emst_get_trace(c, rows.data(), &k);
        last = rows.back();
        iteration = last.iteration;
        if (iteration == 1) {
            // the round-1 cone search is part of the first find-next-closest
            int32_t nt = 4;
            emst_get_times(c, build, &nt);
            last.time_find_next_closest += build[0] + build[1];
        }
        st.cum[0] += last.time_find_next_closest;
        st.cum[1] += last.time_find_min_pair;
        st.cum[2] += last.time_connect_union;
        st.cum[3] += last.time_flatten;
        st.cum[4] += last.time_termination;
        if (trace) st.append(iteration, last);
    }
    // after run(): evaluate, final row to the file and to the console, solution files
    double w = 0;
    int64_t m = 0;
    if ((rc = emst_get_weight(c, &w, &m))) return fail(c, "weight", rc);
    last.size = (int32_t)m;
    last.length = w;
    st.append(iteration, last);
    st.header(std::cout, false);
    st.row(std::cout, iteration, last);
    std::cout << "BEGIN WRITE" << fileSolution << std::endl;
    if ((rc = emst_write_solution(c, fileSolution.c_str()))) {
        emst_destroy(c);
        std::cerr << "write solution failed (" << rc << ")" << std::endl;
        return 1;
    }
    std::cout << "END WRITE" << std::endl;
    emst_destroy(c);
    return 0;
}
