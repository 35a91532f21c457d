// bsar_serve -- serving driver over the reference's stage-boundary file set.
// It loads build/output/proc/ once (the inputs of proc_test_script.cpp:
// 184-440, as bsar_proc does), then images a list of jobs through one
// bsar_server: continuous batching of back-projections over the resident
// dataset, which the reference, imaging one grid per process run, has no
// counterpart for.  Job lines (from --jobs FILE, or standard input):
//
//     <tag> <ix0> <ix1> <iy0> <iy1> <iz0> <iz1> <pulses> <fp64|mixed>
//
// select the sub-grid [ix0, ix1) x [iy0, iy1) x [iz0, iz1) of the file set's
// grid axes and its first <pulses> antenna positions (0 = all); blank lines
// and lines starting with '#' are skipped.  Job <tag> writes
// <dir>/<tag>_3D_matrix.bin and <dir>/<tag>_3D_matrix_grid_{x,y,z}.csv in
// export3DComplexMatrix's formats (data_export.cpp:180-309).  All jobs are
// submitted at once and complete in the server's order (least remaining work
// first); a job's image equals bsar_proc's over the same sub-grid and pulses.
// Exit status: 0 when every job completed, 1 on a load or job error, 2 on a
// malformed command line or job list.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_facade.hpp"
#include "proc_inputs.hpp"

namespace fa = bsar::facade;
using namespace bsar_tools;

namespace {

void usage()
{
    std::printf(
        "usage: bsar_serve [--dir DIR] [--jobs FILE] [--fc HZ] [--fs HZ] [--c M_S] [--factor K] [--device N]\n"
        "                  [--round-ms MS] [--max-jobs N] [--hbm-cap-gb GB] [--check]\n"
        "  job lines: <tag> <ix0> <ix1> <iy0> <iy1> <iz0> <iz1> <pulses> <fp64|mixed>\n"
        "  DIR defaults to build/output/proc; jobs are read from standard input without --jobs.\n"
        "  --check parses the inputs and the job list and prints them without using the GPU.\n");
}

struct Job {
    std::string tag;
    int ix0, ix1, iy0, iy1, iz0, iz1, pulses, precision;
    std::vector<double> gx, gy, gz;
    std::vector<std::complex<double>> img;
    int64_t id = -1;
};

bool valid_tag(const std::string& t)
{
    if (t.empty()) return false;
    for (char ch : t)
        if (!(std::isalnum((unsigned char)ch) || ch == '_' || ch == '-')) return false;
    return true;
}

// Parses the job list against the grid and antenna counts; "" when usable.
std::string read_jobs(std::istream& is, const ProcInputs& in, std::vector<Job>& jobs)
{
    std::string line;
    int lineno = 0;
    std::map<std::string, int> seen;
    while (std::getline(is, line)) {
        ++lineno;
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream ls(line);
        Job j;
        std::string prec, extra;
        if (!(ls >> j.tag >> j.ix0 >> j.ix1 >> j.iy0 >> j.iy1 >> j.iz0 >> j.iz1 >> j.pulses >> prec) || (ls >> extra))
            return "job line " + std::to_string(lineno) + ": expected 9 fields";
        if (!valid_tag(j.tag)) return "job line " + std::to_string(lineno) + ": tag must be [A-Za-z0-9_-]+";
        if (seen.count(j.tag)) return "job line " + std::to_string(lineno) + ": duplicate tag " + j.tag;
        seen[j.tag] = lineno;
        auto range_ok = [](int a, int b, size_t n) { return a >= 0 && a < b && (size_t)b <= n; };
        if (!range_ok(j.ix0, j.ix1, in.gx.size()) || !range_ok(j.iy0, j.iy1, in.gy.size()) ||
            !range_ok(j.iz0, j.iz1, in.gz.size()))
            return "job line " + std::to_string(lineno) + ": index range outside the grid";
        if (j.pulses < 0 || (size_t)j.pulses > in.Tx.size())
            return "job line " + std::to_string(lineno) + ": pulses outside [0, positions]";
        if (prec == "fp64") j.precision = BSAR_FP64;
        else if (prec == "mixed") j.precision = BSAR_MIXED;
        else return "job line " + std::to_string(lineno) + ": precision must be fp64 or mixed";
        j.gx.assign(in.gx.begin() + j.ix0, in.gx.begin() + j.ix1);
        j.gy.assign(in.gy.begin() + j.iy0, in.gy.begin() + j.iy1);
        j.gz.assign(in.gz.begin() + j.iz0, in.gz.begin() + j.iz1);
        jobs.push_back(std::move(j));
    }
    if (jobs.empty()) return "no jobs";
    return "";
}

double quantile(std::vector<double> v, double q)
{
    std::sort(v.begin(), v.end());
    const size_t k = (size_t)std::min<double>((double)v.size() - 1, std::max(0.0, q * (double)v.size() - 0.5));
    return v[k];
}

}  // namespace

int main(int argc, char** argv)
{
    std::string dir = "build/output/proc", jobs_path;
    double fc = 10e9, fs = 60e6, c = 3e8, round_ms = 0.0, cap_gb = 0.0;
    int factor = 4, max_jobs = 0;
    bool check_only = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", what); std::exit(2); }
            return argv[++i];
        };
        if (a == "--dir") dir = next("--dir");
        else if (a == "--jobs") jobs_path = next("--jobs");
        else if (a == "--fc") fc = option_number(next("--fc"), "--fc");
        else if (a == "--fs") fs = option_number(next("--fs"), "--fs");
        else if (a == "--c") c = option_number(next("--c"), "--c");
        else if (a == "--factor") factor = option_integer(next("--factor"), "--factor");
        else if (a == "--device") setenv("BSAR_DEVICE", next("--device").c_str(), 1);
        else if (a == "--round-ms") round_ms = option_number(next("--round-ms"), "--round-ms");
        else if (a == "--max-jobs") max_jobs = option_integer(next("--max-jobs"), "--max-jobs");
        else if (a == "--hbm-cap-gb") cap_gb = option_number(next("--hbm-cap-gb"), "--hbm-cap-gb");
        else if (a == "--check") check_only = true;
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    try {
        ProcInputs in;
        const std::string bad = read_proc_inputs(dir, in);
        if (!bad.empty()) {
            std::fprintf(stderr, "[ERROR] %s\n", bad.c_str());
            return 1;
        }
        print_proc_inputs(in);
        std::vector<Job> jobs;
        std::string jbad;
        if (jobs_path.empty()) {
            jbad = read_jobs(std::cin, in, jobs);
        } else {
            std::ifstream jf(jobs_path);
            if (!jf.is_open()) { std::fprintf(stderr, "[ERROR] cannot open %s\n", jobs_path.c_str()); return 2; }
            jbad = read_jobs(jf, in, jobs);
        }
        if (!jbad.empty()) { std::fprintf(stderr, "[ERROR] %s\n", jbad.c_str()); return 2; }
        std::printf("jobs: %zu\n", jobs.size());
        for (const Job& j : jobs)
            std::printf("job %s: %zu x %zu x %zu voxels x %d pulses, %s\n", j.tag.c_str(), j.gx.size(), j.gy.size(),
                        j.gz.size(), j.pulses ? j.pulses : (int)in.Tx.size(),
                        j.precision == BSAR_MIXED ? "mixed" : "fp64");
        if (check_only) return 0;
        if (in.S < 2 || in.Nd < 1) { std::fprintf(stderr, "[ERROR] rootData needs S >= 2 and >= 1 pulse\n"); return 1; }

        bsar_ctx* ctx = fa::context();
        bsar_dataset* ds = upload_proc_dataset(ctx, in, fc, fs, c, factor);
        const char* dev = std::getenv("BSAR_DEVICE");
        bsar_server_config cfg{};
        cfg.max_jobs = max_jobs;
        cfg.round_seconds = round_ms > 0 ? round_ms * 1e-3 : 0.0;
        cfg.hbm_cap_bytes = cap_gb > 0 ? std::max<int64_t>(1, (int64_t)(cap_gb * 1e9)) : 0;
        bsar_server* srv = nullptr;
        int st = bsar_server_create(dev ? std::atoi(dev) : 0, &cfg, &srv);
        if (st != BSAR_OK) {
            bsar_dataset_destroy(ds);
            std::fprintf(stderr, "[ERROR] server: status %d\n", st);
            return 1;
        }
        std::map<int64_t, Job*> by_id;
        int failed = 0;
        for (Job& j : jobs) {
            j.img.assign(j.gx.size() * j.gy.size() * j.gz.size(), {});
            bsar_job_desc d{};
            d.ds = ds;
            d.pulses = j.pulses;
            d.precision = j.precision;
            d.gx = j.gx.data(); d.nx = (int)j.gx.size();
            d.gy = j.gy.data(); d.ny = (int)j.gy.size();
            d.gz = j.gz.data(); d.nz = (int)j.gz.size();
            d.out = reinterpret_cast<double*>(j.img.data());
            st = bsar_server_submit(srv, &d, &j.id);
            if (st != BSAR_OK) {
                std::fprintf(stderr, "[ERROR] job %s: submit status %d (%s)\n", j.tag.c_str(), st,
                             bsar_server_last_error(srv));
                ++failed;
                continue;
            }
            by_id[j.id] = &j;
        }
        std::vector<double> lat;
        double t_first = 1e300, t_last = 0.0, vp = 0.0;
        for (size_t k = 0; k < by_id.size(); ++k) {
            bsar_job_info info{};
            st = bsar_server_wait_any(srv, -1.0, &info);
            auto it = by_id.find(info.id);
            if (it == by_id.end()) { std::fprintf(stderr, "[ERROR] wait: status %d\n", st); ++failed; break; }
            Job& j = *it->second;
            if (st != BSAR_OK || info.state != BSAR_JOB_DONE) {
                std::fprintf(stderr, "[ERROR] job %s: status %d\n", j.tag.c_str(), info.status);
                ++failed;
                continue;
            }
            const std::string out = dir + "/" + j.tag + "_3D_matrix.bin";
            const int ws = bsar_write_image(out.c_str(), reinterpret_cast<const double*>(j.img.data()),
                                            (int)j.gx.size(), (int)j.gy.size(), (int)j.gz.size(), j.gx.data(),
                                            j.gy.data(), j.gz.data());
            if (ws != BSAR_OK) {
                std::fprintf(stderr, "[ERROR] job %s: cannot write %s (status %d)\n", j.tag.c_str(), out.c_str(), ws);
                ++failed;
                continue;
            }
            const double ms = (info.t_done - info.t_submit) * 1e3;
            lat.push_back(ms);
            t_first = std::min(t_first, info.t_submit);
            t_last = std::max(t_last, info.t_done);
            vp += info.voxel_pulses;
            std::printf("job %s: done in %.1f ms (admitted after %.1f ms), %.3e voxel-pulses, wrote %s\n",
                        j.tag.c_str(), ms, (info.t_admit - info.t_submit) * 1e3, info.voxel_pulses, out.c_str());
        }
        bsar_server_stats ss{};
        bsar_server_get_stats(srv, &ss);
        bsar_server_destroy(srv);
        bsar_dataset_destroy(ds);
        if (!lat.empty()) {
            const double span = t_last - t_first;
            std::printf("[PERFORMANCE] serving: %zu jobs, %.3e voxel-pulses in %.3f s = %.3e voxel-pulse/s aggregate; "
                        "latency p50 %.1f ms, p99 %.1f ms; %lld rounds, device %.3f s\n",
                        lat.size(), vp, span, span > 0 ? vp / span : 0.0, quantile(lat, 0.5), quantile(lat, 0.99),
                        (long long)ss.rounds, ss.device_seconds);
        }
        return failed ? 1 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 1;
    }
}
