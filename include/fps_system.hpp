// fps_system.hpp -- the fTWA system built from the reference's simulation JSON (csrc/system.cpp).
//
// Restates, for the fTWA path only, what FHSystem(Json, folder) (src/FHSystem.cpp:123-326)
// builds: lattice positions from the .xyz file (src/FHLattice.cpp:646-710), distances with the
// periodic minimum image over +-1 cells (:516-539), the hopping tM (src/FHTunnelingInteraction.cpp:
// 16-125, :293-312), U and V (src/FHCoulombInteraction.cpp:13-76, :240-288), the external potential
// (src/FHSystem.cpp:146-169) and the initial occupations (src/FHSystem.cpp:173-293, :536-595).
#pragma once

#include <string>
#include <vector>

#include "fps_json.hpp"

namespace fpsamd {

struct FtwaSystem {
    int L = 0;
    std::vector<double> positions;               // [L][3]
    bool periodic = false;
    double lattice_matrix[9] = {0};              // [3][3]; columns are the lattice vectors
    double same_point_distance = 1e-4;           // include/FHLattice.h:77
    std::vector<double> distance;                // [L][L]
    std::vector<double> tM;                      // [L][L]
    int coulomb_type = 0;                        // 0 null, 1 onSite, 2 general
    std::vector<double> U, V;                    // [L][L]
    std::vector<double> potential;               // [L] or empty (EPnull)
    int pulse_type = 0;                          // 0 none, 1 gaussian, 2 knots
    double pulse_A0 = 0, pulse_t0 = 0, pulse_td = 1, pulse_w0 = 0, pulse_phi = 0;
    double pulse_pol[3] = {0, 0, 0};
    std::vector<double> pulse_knots;
    std::vector<double> displacement;            // [L][L][3] minimum-image r_i - r_j
    std::vector<double> n_up, n_dn;              // [L] initial occupations
    std::vector<double> rot_up, rot_dn;          // [L][L][re,im] initial-state rotations; empty = identity
    std::vector<double> tb_rot;                  // [L][L] tight-binding eigenvectors (columns), when computed
    std::vector<double> tb_energy;               // [L] their eigenvalues, ascending
    std::vector<double> mf_rot_up, mf_rot_dn;    // [L][L] mean-field eigenvectors per spin, when computed
    std::vector<double> mf_energy_up, mf_energy_dn;  // [L] their eigenvalues (last iteration), ascending
    int core_hole_site = -1;                     // lattice site of the core hole (the hole is site L-1), or -1
    int part_up = 0, part_dn = 0;                // upSpinParticleNumber / downSpinParticleNumber (ground-state bases)
    int mf_iterations = 0;                       // self-consistency iterations of the mean-field state
    double mf_residue = 0;                       // its final residue
    std::string ic_type;                         // as read (for messages)
    std::vector<std::string> warnings;           // reference behaviours the build reproduces or deviates from
};

// Builds the system from the "system" object of a simulation file.  folder is the directory of
// the simulation file (with trailing '/'), prefixed to positionsFile as the reference does.
// Returns false with err set on unreadable input or an unsupported request.
bool buildSystem(const fpsjson::Value& systemJson, const std::string& folder, FtwaSystem& out, std::string* err);

// Adds the core hole of a "coreHole" object (src/FHSystem.cpp:295-323) to a built system: L grows by one.
bool addCoreHole(const fpsjson::Value& coreHoleJson, FtwaSystem& s, std::string* err);

// The tight-binding (mean_field false: src/FHSystem.cpp:909-937) or mean-field (true: :940-1019)
// eigenbasis of the built system for n_up / n_dn particles, into tb_rot or mf_rot_up/mf_rot_dn.
bool computeGroundState(FtwaSystem& s, bool mean_field, int n_up, int n_dn, std::string* err);

// Minimum-image distance between two sites (src/FHLattice.cpp:516-539).
double siteDistance(const FtwaSystem& s, int i, int j, double* disp3);

// The "tbGroundState" (src/FHSystem.cpp:1143-1168) and "mfGroundState" (:1171-1230) blocks of a system or
// results file: eigenvalues ascending, rotation matrices [L][L] row-major (eigenvectors in the columns).
fpsjson::Value tbGroundStateJson(const std::vector<double>& energy, const std::vector<double>& rot);
fpsjson::Value mfGroundStateJson(const std::vector<double>& eu, const std::vector<double>& ed,
                                 const std::vector<double>& ru, const std::vector<double>& rd, int iterations,
                                 double residue);

// ---- the system generator of makeSitesSystem (csrc/lattice_gen.cpp; the fps_make_system CLI) ----

// A generated lattice: its type name as the system file writes it ("grid", "bcc", "fcc", "triangular",
// "honeyCombs", "graphene", "moireHC", "moire"; "kagome" and unknown names ("none") keep the reference's
// two-site default), positions [L][3] and the box, row-major [3][3] with the lattice vectors in columns.
struct GeneratedLattice {
    std::string type;
    std::vector<double> positions;
    double matrix[9] = {0};
};

// FHLattice(std::string, std::vector<double>) (src/FHLattice.cpp:169-321): parameters padded to five
// with 1 (makeSitesSystem.cpp:128-131) and read per type as
//   grid, bcc, fcc, triangular: nx ny nz a;  honeyCombs: nCombs a;  graphene: nx ny a;
//   moireHC: nCombs a b theta;  moire: nx ny a b theta (theta in degrees, b the layer spacing).
bool generateLattice(const std::string& typeName, const std::vector<double>& parameters, GeneratedLattice& out);

// The .xyz file FHLattice::writePositions writes (6 significant digits, src/FHLattice.cpp:715-738).
bool writeXyzFile(const std::string& path, const std::vector<double>& positions, const std::string& siteName);

// The simulation file makeSitesSystem writes (makeSitesSystem.cpp:170-205): method "HF", its default
// parameters and results, and a system of nearest-neighbour hopping -1 at distance 1, on-site U = 1, no
// external potential, with the initial condition ic ("spinWave", "chargeWave", "tbGroundState" or
// "mfGroundState").  The ground-state blocks of the last two are added by the caller.
fpsjson::Value sitesSystemJson(const std::string& name, const GeneratedLattice& lat, const std::string& ic);

}  // namespace fpsamd
