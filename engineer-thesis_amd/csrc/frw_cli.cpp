// Command-line driver with the reference's flags (randomWalk/main.cpp:15-140):
//   -S/--source PATH  -I/--iterations N  -O/--object ID  -G/--GPU
//   -M/--measure      -L/--layer K       --help
// Defaults as randomWalk/defines.h:5-7,26-28.  The walk seed comes from the
// FRW_SEED environment variable (default 1).
#include <cstdio>
#include <fstream>
#include <string>

#include "randomwalk.h"

static const char* kHelp =
    "\t-S\t--source\t\tlayout file\n"
    "\t-M\t--measure\t\tprint build/walk timings\n"
    "\t-O\t--object\t\tmaster rect id\n"
    "\t-I\t--iterations\t\tnumber of walks\n"
    "\t-G\t--GPU\t\t\texact-search engine (default: reference-compatible tree and queries)\n"
    "\t-L\t--layer\t\t\tlayer number\n"
    "\t\t--help\t\t\tthis text\n";

static bool read_int(int argc, char** argv, int& i, int& out, const char* flag)
{
    if (i + 1 >= argc) {
        std::printf("%s needs one argument.\n", flag);
        return false;
    }
    try {
        out = std::stoi(std::string(argv[++i]));
    } catch (...) {
        std::printf("Not a number: %s\n", argv[i]);
        return false;
    }
    return true;
}

int main(int argc, char* argv[])
{
    std::string path = "../tests/test";
    bool gpu = false, measure = false;
    int rect = 10, iters = 1000, layer = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string o(argv[i]);
        if (o == "--help") {
            std::printf("%s", kHelp);
            if (argc == 2) return 0;
        } else if (o == "-G" || o == "--GPU") {
            gpu = true;
        } else if (o == "-M" || o == "--measure") {
            measure = true;
        } else if (o == "-S" || o == "--source") {
            if (i + 1 >= argc) {
                std::printf("--source needs one argument.\n");
                return -1;
            }
            path = argv[++i];
        } else if (o == "-I" || o == "--iterations") {
            if (!read_int(argc, argv, i, iters, "--iterations")) return -1;
        } else if (o == "-O" || o == "--object") {
            if (!read_int(argc, argv, i, rect, "--object")) return -1;
        } else if (o == "-L" || o == "--layer") {
            if (!read_int(argc, argv, i, layer, "--layer")) return -1;
        } else {
            std::printf("%sUnknown argument: %s\n", kHelp, o.c_str());
            return -1;
        }
    }
    if (!initCuda(argc, argv)) return 1;
    if (!std::ifstream(path).good()) {
        std::fprintf(stderr, "No such file!\n");
        return 1;
    }
    runRandomWalk(&path[0], iters, rect, gpu, measure, layer);
    return 0;
}
