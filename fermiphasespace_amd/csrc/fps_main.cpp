// fps_main.cpp -- command-line driver: `fps_mi355x <sim.json> [out.json]`.
//
// The fTWA branch of the reference driver (fps.cpp:88-207, :264-270): read the simulation file,
// build the system, run, write <folder><sim_name>_fTWA_res.json, print the computation time.
#include <chrono>
#include <cstdio>
#include <iostream>

#include "../../include/fps_ftwa.h"

int main(int argc, char** argv)
{
    std::cout << "--- Fermi-Hubbard Dynamic with Phase-space methods (fTWA, MI355X engine) --- " << std::endl;
    int ma = 0, mi = 0, pa = 0;
    fpsftwa_version(&ma, &mi, &pa);
    std::cout << "== engine version:  " << ma << "." << mi << "." << pa << std::endl;
    if (argc < 2) {
        std::cout << "Need a json file. " << std::endl;
        return 2;
    }
    char dev[256] = {0};
    int cus = 0;
    double gib = 0;
    if (fpsftwa_device_info(dev, sizeof dev, &cus, &gib) == FPSFTWA_OK)
        std::cout << "== device: " << dev << ", " << cus << " CUs, " << gib << " GiB" << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = fpsftwa_run_json_file(argv[1], argc > 2 ? argv[2] : nullptr);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != FPSFTWA_OK) {
        std::cerr << "error (" << rc << "): " << fpsftwa_last_error() << std::endl;
        return 1;
    }
    const std::chrono::duration<double> el = t1 - t0;
    std::cout << "== Computation time : " << el.count() << " s." << std::endl;
    std::cout << "== Program closing." << std::endl;
    return 0;
}
