// wurst_main.cpp -- command line of the MI355X `nextevent_path_tracer`, with
// the reference's options (src/main.cpp:5-48):
//   wurst_amd --simple2parse scene.json
//   wurst_amd --jsimple2parse '{"camera": ..., "meshes": [...], "render": [...]}'
//   wurst_amd                      (reads option/value pairs from default.txt)
// Options of this build:
//   --check              parse and validate the scene, render nothing (no GPU)
//   --dump DIR           write the flattened scene arrays into DIR
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "wurst_host.hpp"

int main(int argc, const char *argv[]) {
    wurst::ParseOptions opt;
    std::vector<std::pair<std::string, std::string>> jobs;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--check") { opt.check_only = true; continue; }
        if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); return 2; }
        const std::string v = argv[++i];
        if (a == "--dump") opt.dump_dir = v;
        else if (a == "--simple2parse" || a == "--jsimple2parse") jobs.emplace_back(a, v);
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    if (jobs.empty()) {
        std::ifstream ifs("default.txt");
        if (!ifs.is_open()) { std::ofstream created("default.txt"); return 0; }
        for (std::string line; std::getline(ifs, line);) {
            std::stringstream ss(line);
            std::string option, info;
            ss >> option >> std::quoted(info);
            if (!option.empty() && option[0] != '#' && !info.empty()) jobs.emplace_back(option, info);
        }
    }
    try {
        for (const auto &job : jobs) {
            if (job.first == "--simple2parse") wurst::Simple2Parser::Parse(job.second, opt);
            else if (job.first == "--jsimple2parse") wurst::Simple2Parser::Parse(wj::parse(job.second), ".", opt);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
