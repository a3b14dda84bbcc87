// make_system_main.cpp -- `fps_make_system <name> <type> <param1> ... [ic_sw|ic_cw|ic_tbgs|ic_mfgs]`.
//
// The system generator of the reference (makeSitesSystem.cpp:47-255): builds a lattice of the given type,
// writes systemsFolder/<name>.xyz and systemsFolder/<name>.json (an "HF" simulation file of that system),
// with the tight-binding or mean-field ground state written into the system when the initial condition
// asks for it.  Arguments after the type: numbers are the geometry parameters (after the first word that
// is not a number, later numbers are not read, as in the reference), ic_* words the initial condition.
// Unlike the reference, a missing systemsFolder is created.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/fps_system.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cout << "To make a new system :" << std::endl;
        std::cout << "./makeSitesSystem <nameOfTheSystem> <geometryOfTheSites> <param1> <param2> ..." << std::endl;
        return 0;
    }
    std::cout << "== Writing a new FH system." << std::endl;
    const std::string folder = "systemsFolder";
    std::string name = argv[1], type = "grid", ic = "spinWave";
    std::vector<double> geom;
    bool numbers = true;
    for (int k = 2; k < argc; ++k) {
        const std::string a = argv[k];
        if (k == 2) {
            type = a;
        } else if (a == "ic_sw") {
            ic = "spinWave";
        } else if (a == "ic_cw") {
            ic = "chargeWave";
        } else if (a == "ic_tbgs") {
            ic = "tbGroundState";
        } else if (a == "ic_mfgs") {
            ic = "mfGroundState";
        } else {
            char* end = nullptr;
            const double v = std::strtod(a.c_str(), &end);
            if (end == a.c_str()) numbers = false;  // not a number: this and later words are not read
            if (numbers) geom.push_back(v);
        }
    }
    fpsamd::GeneratedLattice lat;
    fpsamd::generateLattice(type, geom, lat);
    const int L = static_cast<int>(lat.positions.size() / 3);
    mkdir(folder.c_str(), 0755);
    const std::string xyz = folder + "/" + name + ".xyz", js = folder + "/" + name + ".json";
    if (!fpsamd::writeXyzFile(xyz, lat.positions, "C")) {
        std::cerr << "error: cannot write " << xyz << std::endl;
        return 1;
    }
    fpsjson::Value sim = fpsamd::sitesSystemJson(name, lat, ic);
    // particle numbers: spin wave ceil(L/2) up on even sites and floor(L/2) down; charge wave ceil(L/2) of
    // each on even sites; ground states ceil(L/2) up and floor(L/2) down (src/FHSystem.cpp:453-533)
    const int up = (L + 1) / 2, dn = ic == "chargeWave" ? (L + 1) / 2 : L / 2;
    if (ic == "tbGroundState" || ic == "mfGroundState") {
        fpsamd::FtwaSystem s;
        std::string err;
        if (!fpsamd::buildSystem(sim["system"], folder, s, &err)) {
            std::cerr << "error: " << err << std::endl;
            return 1;
        }
        if (ic == "tbGroundState")
            sim["system"]["tbGroundState"] = fpsamd::tbGroundStateJson(s.tb_energy, s.tb_rot);
        else
            sim["system"]["mfGroundState"] = fpsamd::mfGroundStateJson(s.mf_energy_up, s.mf_energy_dn, s.mf_rot_up,
                                                                       s.mf_rot_dn, s.mf_iterations, s.mf_residue);
    }
    {
        std::ofstream f(js);
        if (!f.is_open()) {
            std::cerr << "error: cannot write " << js << std::endl;
            return 1;
        }
        f << fpsjson::write(sim);
    }
    std::cout << "== " << name << " system created with " << L << " sites and " << up + dn << " particles." << std::endl;
    std::cout << "== Particle positions in " << folder << "/" << name << ".xyz" << std::endl;
    return 0;
}
