// lulesh_main.cc -- command-line driver of the MI355X LULESH step.
//
// The C++ caller of the C ABI (include/lulesh_mi355x.h).  Options follow
// lulesh-2.0.3 (reference/C/lulesh-2.0.3/lulesh-util.cc:28-168):
//   -s N   elements per edge (default 30)      -i N   maximum cycles
//   -r N   regions (default 11)                -b N   region balance (1)
//   -c N   region cost (1)                     -p     print every cycle
//   -q     quiet                               -h     help
// plus  --ic 1|2  initial conditions of LULESH 1.0 (LULESH-OMP.cc) or 2.0.3
// (default 2).  The final report is the reference's verification block
// (lulesh-util.cc:172-219; LULESH-OMP.cc:3195-3223) with the elapsed time of
// the cycle loop measured around the library calls.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "lulesh_mi355x.h"

namespace {

void usage(const char *prog)
{
    std::printf(
        "Usage: %s [opts]\n"
        " where [opts] is one or more of:\n"
        " -q              : quiet mode - suppress all stdout\n"
        " -i <iterations> : number of cycles to run\n"
        " -s <size>       : length of cube mesh along side\n"
        " -r <numregions> : Number of distinct regions (def: 11)\n"
        " -b <balance>    : Load balance between regions of a domain (def: 1)\n"
        " -c <cost>       : Extra cost of more expensive regions (def: 1)\n"
        " -f <numfiles>   : Number of files to split viz dump into (def: 1; one\n"
        "                   VTK file is written whatever the count)\n"
        " -p              : Print out progress\n"
        " -v              : Output viz file: lulesh_plot_c<cycle>.vtk\n"
        " --ic <1|2>      : LULESH 1.0 or 2.0 initial conditions (def: 2)\n"
        " -u <file>       : unstructured mesh from a .lmesh file (LULESH 1.0)\n"
        " --vtk <file>    : write the final state as a VTK file\n"
        " -h              : This message\n",
        prog);
}

bool int_arg(int argc, char **argv, int &i, int &out)
{
    if (i + 1 >= argc) {
        std::fprintf(stderr, "missing integer argument to %s\n", argv[i]);
        return false;
    }
    char *end = nullptr;
    const long v = std::strtol(argv[i + 1], &end, 10);
    if (!end || *end != '\0') {
        std::fprintf(stderr, "bad integer argument to %s: %s\n", argv[i], argv[i + 1]);
        return false;
    }
    out = static_cast<int>(v);
    ++i;
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    lulesh_config cfg;
    lulesh_config_default(&cfg);
    bool quiet = false, progress = false, viz = false;
    int num_files = 1;
    std::string lmesh, vtk;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-s") {
            if (!int_arg(argc, argv, i, cfg.nx)) return 1;
        } else if (a == "-i") {
            if (!int_arg(argc, argv, i, cfg.max_cycles)) return 1;
        } else if (a == "-r") {
            if (!int_arg(argc, argv, i, cfg.num_reg)) return 1;
        } else if (a == "-b") {
            if (!int_arg(argc, argv, i, cfg.balance)) return 1;
        } else if (a == "-c") {
            if (!int_arg(argc, argv, i, cfg.cost)) return 1;
        } else if (a == "--ic") {
            if (!int_arg(argc, argv, i, cfg.ic_mode)) return 1;
        } else if (a == "-u" || a == "--vtk") {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing file argument to %s\n", argv[i]);
                return 1;
            }
            (a == "-u" ? lmesh : vtk) = argv[++i];
        } else if (a == "-f") {
            if (!int_arg(argc, argv, i, num_files)) return 1;
            if (num_files < 1) {
                std::fprintf(stderr, "-f needs a positive number of files\n");
                return 1;
            }
        } else if (a == "-v") {
            viz = true;
        } else if (a == "-p") {
            progress = true;
        } else if (a == "-q") {
            quiet = true;
        } else if (a == "-h") {
            usage(argv[0]);
            return 0;
        } else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            usage(argv[0]);
            return 1;
        }
    }

    lulesh_domain *d = nullptr;
    int rc;
    if (!lmesh.empty()) {
        cfg.ic_mode = LULESH_IC_1_0;
        rc = lulesh_create_lmesh(lmesh.c_str(), &cfg, &d);
    } else {
        rc = lulesh_create(&cfg, &d);
    }
    if (rc != LULESH_OK) {
        std::fprintf(stderr, "creating the domain failed (%d)\n", rc);
        return 1;
    }
    lulesh_stats st0;
    lulesh_get_stats(d, &st0);
    if (!quiet) {
        if (lmesh.empty()) {
            std::printf("Running problem size %d^3 per domain until completion\n", cfg.nx);
        } else {
            std::printf("Running mesh %s until completion\n", lmesh.c_str());
        }
        std::printf("Num processors: 1 (MI355X)\n");
        std::printf("Initial conditions: LULESH %s\n", cfg.ic_mode == 1 ? "1.0" : "2.0");
        std::printf("Total number of elements: %d\n\n", st0.num_elem);
    }

    const auto t0 = std::chrono::steady_clock::now();
    int total = 0;
    if (progress) {
        for (;;) {
            int n = 0;
            rc = lulesh_step(d, 1, &n);
            if (rc != LULESH_OK || n == 0) break;
            total += n;
            lulesh_stats st;
            lulesh_get_stats(d, &st);
            if (!quiet) {
                std::printf("cycle = %d, time = %e, dt=%e\n", st.cycle, st.time,
                            st.deltatime);
            }
        }
    } else {
        rc = lulesh_run(d, &total);
    }
    const auto t1 = std::chrono::steady_clock::now();
    const double elapsed = std::chrono::duration<double>(t1 - t0).count();
    if (rc != LULESH_OK) {
        std::fprintf(stderr, "step failed (%d): %s\n", rc, lulesh_last_error(d));
        lulesh_destroy(d);
        return rc == LULESH_VOLUME_ERROR || rc == LULESH_QSTOP_ERROR ? -rc : 1;
    }

    lulesh_stats st;
    lulesh_get_stats(d, &st);
    double mabs = 0, tabs = 0, mrel = 0;
    lulesh_symmetry(d, &mabs, &tabs, &mrel);  // the cube only
    int status = 0;
    if (viz && vtk.empty()) {
        // the reference's dump name (lulesh-2.0.3/lulesh-viz.cc:62)
        vtk = "lulesh_plot_c" + std::to_string(st.cycle) + ".vtk";
    }
    if (!vtk.empty() && lulesh_write_domain(d, vtk.c_str()) != LULESH_OK) {
        std::fprintf(stderr, "writing %s failed\n", vtk.c_str());
        status = 1;
    }
    if (!quiet) {
        const double zones = double(st.num_elem);
        const double grind = (elapsed * 1e6 / st.cycle) / zones;
        std::printf("Run completed:  \n");
        std::printf("   Problem size        =  %i \n", cfg.nx);
        std::printf("   MPI tasks           =  1 \n");
        std::printf("   Iteration count     =  %i \n", st.cycle);
        std::printf("   Final Origin Energy = %12.6e \n", st.e_origin);
        std::printf("   Testing Plane 0 of Energy Array on rank 0:\n");
        std::printf("        MaxAbsDiff   = %12.6e\n", mabs);
        std::printf("        TotalAbsDiff = %12.6e\n", tabs);
        std::printf("        MaxRelDiff   = %12.6e\n\n", mrel);
        std::printf("\nElapsed time         = %10.2f (s)\n", elapsed);
        std::printf("Grind time (us/z/c)  = %10.8g (per dom)  (%10.8g overall)\n",
                    grind, grind);
        std::printf("FOM                  = %10.8g (z/s)\n", 1000.0 / grind);
        std::printf("zones*cycles/s       = %10.6e\n\n", zones * st.cycle / elapsed);
    }
    lulesh_destroy(d);
    return status;
}
