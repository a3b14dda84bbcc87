// system.cpp -- JSON -> fTWA system (see system.hpp for the reference functions restated).
#include "../../include/fps_system.hpp"

#include "eigen.hpp"

#include <cmath>
#include <fstream>
#include <sstream>

namespace fpsamd {

namespace {

bool readXyz(const std::string& path, std::vector<double>& pos, int& L, std::string* err)
{
    // .xyz: site count, comment line, then "name x [y [z]]" per site (src/FHLattice.cpp:646-710)
    std::ifstream f(path);
    if (!f.good()) {
        if (err) *err = "cannot open positions file " + path;
        return false;
    }
    std::string line;
    if (!std::getline(f, line)) {
        if (err) *err = "empty positions file " + path;
        return false;
    }
    int n = std::atoi(line.c_str());
    std::getline(f, line);  // comment
    pos.clear();
    int found = 0;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string name;
        if (!(ss >> name)) continue;  // blank line
        double c[3] = {0, 0, 0};
        std::string tok;
        int k = 0;
        while (ss >> tok) {
            if (k < 3) c[k] = std::strtod(tok.c_str(), nullptr);
            ++k;
        }
        pos.push_back(c[0]);
        pos.push_back(c[1]);
        pos.push_back(c[2]);
        ++found;
    }
    if (found != n) {
        if (err) *err = "positions file " + path + ": " + std::to_string(n) + " sites announced, " +
                        std::to_string(found) + " found";
        return false;
    }
    L = n;
    return true;
}

inline double sq(double x) { return x * x; }

}  // namespace

double siteDistance(const FtwaSystem& s, int i, int j, double* disp3)
{
    const double* a = &s.positions[3 * i];
    const double* b = &s.positions[3 * j];
    double d0[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double best[3] = {d0[0], d0[1], d0[2]};
    double dmin = std::sqrt(sq(d0[0]) + sq(d0[1]) + sq(d0[2]));
    if (s.periodic) {
        const double* m = s.lattice_matrix;  // m[r*3+c]; column c is lattice vector c
        for (int z = -1; z < 2; ++z)
            for (int y = -1; y < 2; ++y)
                for (int x = -1; x < 2; ++x) {
                    double d[3];
                    for (int r = 0; r < 3; ++r) d[r] = d0[r] + x * m[r * 3 + 0] + y * m[r * 3 + 1] + z * m[r * 3 + 2];
                    double n = std::sqrt(sq(d[0]) + sq(d[1]) + sq(d[2]));
                    if (n < dmin) {
                        dmin = n;
                        best[0] = d[0]; best[1] = d[1]; best[2] = d[2];
                    }
                }
    }
    if (disp3) { disp3[0] = best[0]; disp3[1] = best[1]; disp3[2] = best[2]; }
    return dmin;
}

bool addCoreHole(const fpsjson::Value& ch, FtwaSystem& s, std::string* err)
{
    using fpsjson::Value;
    // FHCoreHole(Json) defaults (src/FHCoreHole.cpp:10-26) and the extension of FHSystem(Json)
    // (src/FHSystem.cpp:295-323): one more site L, coupled to `site` by the hopping X, on-site
    // potential eps, interactions addSite([.. G at site .., Uc at L]) (src/FHCoulombInteraction.cpp:
    // 189-234), occupations IC = [up, down] (default [1, 0]).
    const int L = s.L;
    const int64_t site = ch.get("site", Value((int64_t)0)).asInt();
    const double eps = ch.get("eps", Value(-2.0)).asDouble();
    const double X = ch.get("X", Value(-1.0)).asDouble();
    const double Uc = ch.get("Uc", Value(1.0)).asDouble();
    const double G = ch.get("G", Value(1.0)).asDouble();
    double occ_up = 1.0, occ_dn = 0.0;
    const Value& ic = ch["IC"];
    if (ic.isArray() && ic.size() > 0) occ_up = ic[0].asDouble();
    if (ic.isArray() && ic.size() > 1) occ_dn = ic[1].asDouble();
    if (site < 0 || site >= L) {
        if (err) *err = "coreHole site " + std::to_string(site) + " outside the lattice";
        return false;
    }
    if (s.coulomb_type == 2) {
        // for CEdistanceInteraction the reference grows U and V to L+2 while every other matrix is
        // L+1 (src/FHCoulombInteraction.cpp:213-222): no consistent system exists
        if (err) *err = "a core hole with a distance Coulomb interaction has no consistent form in the reference";
        return false;
    }
    const int L1 = L + 1;
    auto grow = [&](std::vector<double>& m) {
        std::vector<double> g((size_t)L1 * L1, 0.0);
        if (!m.empty())
            for (int i = 0; i < L; ++i)
                for (int j = 0; j < L; ++j) g[(size_t)i * L1 + j] = m[(size_t)i * L + j];
        m.swap(g);
    };
    grow(s.tM);
    s.tM[(size_t)site * L1 + L] = X;
    s.tM[(size_t)L * L1 + site] = X;
    grow(s.U);
    s.V.assign((size_t)L1 * L1, 0.0);  // null / onSite: V starts from zero
    s.U[(size_t)site * L1 + L] += G / 2;
    s.U[(size_t)L * L1 + site] += G / 2;
    s.U[(size_t)L * L1 + L] += Uc;  // Uc/2 from the column and Uc/2 from the row
    s.V[(size_t)site * L1 + L] = G / 2;
    s.V[(size_t)L * L1 + site] = G / 2;
    s.coulomb_type = 2;  // CEgiven: the dynamics adds the same-spin V term
    // the reference rebuilds the external potential from scratch before adding the hole's
    // (src/FHSystem.cpp:311-312): a constant potential or pulse given in the file is dropped
    if (!s.potential.empty() || s.pulse_type != 0)
        s.warnings.push_back("coreHole: the external potential of the file is replaced by the hole potential alone "
                             "(as in the reference)");
    s.pulse_type = 0;
    if (eps == 0.0) {
        s.potential.clear();
    } else {
        s.potential.assign(L1, 0.0);
        s.potential[L] = eps;
    }
    // geometry: the hole sits at its site (distances and displacements of the hole are the site's)
    std::vector<double> pos = s.positions, dist((size_t)L1 * L1, 0.0), disp((size_t)L1 * L1 * 3, 0.0);
    pos.insert(pos.end(), s.positions.begin() + 3 * site, s.positions.begin() + 3 * site + 3);
    for (int i = 0; i < L1; ++i)
        for (int j = 0; j < L1; ++j) {
            const int a = i == L ? (int)site : i, b = j == L ? (int)site : j;
            dist[(size_t)i * L1 + j] = s.distance[(size_t)a * L + b];
            for (int k = 0; k < 3; ++k) disp[((size_t)i * L1 + j) * 3 + k] = s.displacement[((size_t)a * L + b) * 3 + k];
        }
    s.positions.swap(pos);
    s.distance.swap(dist);
    s.displacement.swap(disp);
    // initial state (src/FHfTWASolver.cpp:156-184): occupations and rotations gain the hole
    s.n_up.push_back(occ_up);
    s.n_dn.push_back(occ_dn);
    for (std::vector<double>* r : {&s.rot_up, &s.rot_dn}) {
        if (r->empty()) continue;
        std::vector<double> g((size_t)L1 * L1 * 2, 0.0);
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                g[((size_t)i * L1 + j) * 2] = (*r)[((size_t)i * L + j) * 2];
                g[((size_t)i * L1 + j) * 2 + 1] = (*r)[((size_t)i * L + j) * 2 + 1];
            }
        g[((size_t)L * L1 + L) * 2] = 1.0;
        r->swap(g);
    }
    s.core_hole_site = (int)site;
    s.L = L1;
    return true;
}

namespace {

fpsjson::Value vectorJson(const std::vector<double>& v)
{
    fpsjson::Value a = fpsjson::Value::array();
    for (double x : v) a.append(x);
    return a;
}

// [n][n] row-major -> array of rows (FHSystem::writeTBGroundStateJson writes V row by row, real parts)
fpsjson::Value matrixJson(const std::vector<double>& m)
{
    const size_t n = static_cast<size_t>(std::llround(std::sqrt(static_cast<double>(m.size()))));
    fpsjson::Value a = fpsjson::Value::array();
    for (size_t i = 0; i < n; ++i) {
        fpsjson::Value r = fpsjson::Value::array();
        for (size_t j = 0; j < n; ++j) r.append(m[i * n + j]);
        a.append(r);
    }
    return a;
}

}  // namespace

fpsjson::Value tbGroundStateJson(const std::vector<double>& energy, const std::vector<double>& rot)
{
    fpsjson::Value o = fpsjson::Value::object();
    o["energyVector"] = vectorJson(energy);
    o["rotationMatrix"] = matrixJson(rot);
    return o;
}

fpsjson::Value mfGroundStateJson(const std::vector<double>& eu, const std::vector<double>& ed,
                                 const std::vector<double>& ru, const std::vector<double>& rd, int iterations,
                                 double residue)
{
    fpsjson::Value o = fpsjson::Value::object();
    o["energyVectorUp"] = vectorJson(eu);
    o["energyVectorDown"] = vectorJson(ed);
    o["rotationMatrixUp"] = matrixJson(ru);
    o["rotationMatrixDown"] = matrixJson(rd);
    o["iterationsNeeded"] = iterations;
    o["residueComputation"] = residue;
    return o;
}

bool computeGroundState(FtwaSystem& s, bool mean_field, int n_up, int n_dn, std::string* err)
{
    const int L = s.L;
    std::vector<double> ru((size_t)L * L), rd((size_t)L * L), eu(L), ed(L);
    int it = 0;
    double res = 0.0;
    if (ground_state_rotation(L, s.tM.data(), s.U.empty() ? nullptr : s.U.data(), n_up, n_dn, mean_field, ru.data(),
                              rd.data(), &it, &res, eu.data(), ed.data()) != 0) {
        if (err) *err = "eigen-decomposition of the hopping matrix did not converge";
        return false;
    }
    if (!mean_field) {
        s.tb_rot = std::move(ru);
        s.tb_energy = std::move(eu);
        return true;
    }
    s.mf_rot_up = std::move(ru);
    s.mf_rot_dn = std::move(rd);
    s.mf_energy_up = std::move(eu);
    s.mf_energy_dn = std::move(ed);
    s.mf_iterations = it;
    s.mf_residue = res;
    if (it >= 100)
        s.warnings.push_back("the mean-field ground state did not converge in 100 iterations (residue " +
                             std::to_string(res) + "); its last iterate is used, as in the reference");
    return true;
}

bool buildSystem(const fpsjson::Value& sys, const std::string& folder, FtwaSystem& out, std::string* err)
{
    using fpsjson::Value;
    FtwaSystem s;
    // ---- lattice (src/FHLattice.cpp:101-167) ----
    const Value& lat = sys["lattice"];
    std::string posFile = lat.get("positionsFile", Value("")).asString();
    if (!folder.empty()) posFile = folder + "/" + posFile;
    if (!readXyz(posFile, s.positions, s.L, err)) return false;
    if (s.L <= 0) {
        if (err) *err = "lattice has no sites";
        return false;
    }
    s.periodic = lat.get("periodic", Value(false)).asBool();
    const Value& lm = lat["latticeMatrix"];
    if (!lm.isNull())
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) s.lattice_matrix[r * 3 + c] = lm[r][c].asDouble();
    const int L = s.L;
    s.distance.assign((size_t)L * L, 0.0);
    s.displacement.assign((size_t)L * L * 3, 0.0);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
            s.distance[(size_t)i * L + j] = siteDistance(s, i, j, &s.displacement[((size_t)i * L + j) * 3]);
    const double ths = s.same_point_distance;

    // ---- tunnelling (src/FHTunnelingInteraction.cpp:16-125, addTunneling :293-312) ----
    s.tM.assign((size_t)L * L, 0.0);
    auto addTunneling = [&](double dist, double E) {
        for (int i = 0; i < L; ++i)
            for (int j = i + 1; j < L; ++j)
                if (sq(s.distance[(size_t)i * L + j] - dist) < ths) {
                    s.tM[(size_t)i * L + j] += E;
                    s.tM[(size_t)j * L + i] += E;
                }
    };
    const Value& tun = sys["tunnelingInteraction"];
    std::string ttype = tun.get("type", Value("null")).asString();
    if (ttype == "nearestNeighbors") {
        addTunneling(tun.get("neighborDistance", Value(1.0)).asDouble(),
                     tun.get("neighborInteraction", Value(-1.0)).asDouble());
    } else if (ttype == "distanceInteraction") {
        const Value& ds = tun["distances"];
        const Value& es = tun["energies"];
        if (!ds.isNull()) {
            if (ds.size() == es.size()) {
                for (size_t k = 0; k < ds.size(); ++k) addTunneling(ds[k].asDouble(), es[k].asDouble());
            } else {
                s.warnings.push_back("Problem in the tunneling energy parameters.");
            }
        }
    } else if (ttype == "distanceInteraction1" || ttype == "distanceInteraction2" || ttype == "given") {
        s.warnings.push_back("tunnelingInteraction type '" + ttype + "' adds no hopping (as in the reference)");
    } else if (sys.isMember("EnergyT") || sys.isMember("parametersT")) {
        s.warnings.push_back("legacy keys EnergyT/parametersT are not read (as in the reference): zero hopping");
    }

    // ---- Coulomb (src/FHCoulombInteraction.cpp:13-76, :240-288) ----
    s.U.assign((size_t)L * L, 0.0);
    s.V.assign((size_t)L * L, 0.0);
    auto addCoulomb = [&](double dist, double E) {
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j)
                if (sq(s.distance[(size_t)i * L + j] - dist) < ths) {
                    s.U[(size_t)i * L + j] += E;
                    if (i != j) s.V[(size_t)i * L + j] += E;
                }
    };
    const Value& cou = sys["coulombInteraction"];
    std::string ctype = cou.get("type", Value("null")).asString();
    if (ctype == "onSite") {
        s.coulomb_type = 1;
        addCoulomb(0.0, cou.get("onSiteInteraction", Value(1.0)).asDouble());
    } else if (ctype == "distanceInteraction") {
        // JSON form of the distance-table constructor (:57-76), keys as its writer emits them (:160-177)
        s.coulomb_type = 2;
        const Value& ds = cou["distances"];
        const Value& es = cou["energies"];
        size_t n = std::max(ds.size(), es.size());
        for (size_t k = 0; k < n; ++k) addCoulomb(ds[k].asDouble(), es[k].asDouble());
    } else {
        s.coulomb_type = 0;
        if (ctype != "null") s.warnings.push_back("coulombInteraction type '" + ctype + "' is not parsed: no interaction");
    }

    // ---- external potential (src/FHSystem.cpp:146-169) ----
    const Value& ep = sys["externalPotential"];
    if (!ep.isNull()) {
        std::string etype = ep.get("type", Value("null")).asString();
        if (etype == "constant") {
            s.potential.assign(L, 0.0);
            const Value& dg = ep["diagonal"];  // the key the reference writer emits (src/FHExternalPotential.cpp:146-164)
            if (dg.isArray()) {
                for (int i = 0; i < L && i < (int)dg.size(); ++i) s.potential[i] = dg[i].asDouble();
            } else {
                s.warnings.push_back("constant potential without 'diagonal': zero potential");
            }
        } else if (etype == "coreHole") {
            s.potential.assign(L, 0.0);
            int64_t site = ep.get("site", Value(0)).asInt();
            double hp = ep.get("holePotential", Value(-1.0)).asDouble();
            if (site >= 0 && site < L) s.potential[site] = hp;
        } else if (etype == "pulse") {
            // Peierls phase on the hopping (src/FHExternalPotentialPulse.cpp:86-110); keys as its writer (:112-122)
            s.pulse_A0 = ep.get("A0", Value(0.0)).asDouble();
            s.pulse_t0 = ep.get("t0", Value(0.0)).asDouble();
            s.pulse_td = ep.get("td", Value(1.0)).asDouble();
            s.pulse_w0 = ep.get("w0", Value(0.0)).asDouble();
            s.pulse_phi = ep.get("phi", Value(0.0)).asDouble();
            const Value& pol = ep["polarization"];
            if (pol.isArray()) {
                for (int k = 0; k < 3; ++k) s.pulse_pol[k] = pol[k].asDouble();
            } else {  // makePotential_pulse's polarisation (src/FHSystem.cpp:818-828)
                s.pulse_pol[0] = 1.0 / std::sqrt(2.0);
                s.pulse_pol[1] = 1.0 / std::sqrt(2.0);
            }
            const Value& kn = ep["knots"];
            if (kn.isArray() && kn.size() > 0) {
                s.pulse_type = 2;
                for (size_t k = 0; k < kn.size(); ++k) s.pulse_knots.push_back(kn[k].asDouble());
            } else {
                s.pulse_type = (s.pulse_A0 != 0.0) ? 1 : 0;
            }
            const Value& dg = ep["diagonal"];
            if (dg.isArray()) {
                s.potential.assign(L, 0.0);
                for (int i = 0; i < L && i < (int)dg.size(); ++i) s.potential[i] = dg[i].asDouble();
            }
        }
    }

    // ---- initial condition (src/FHSystem.cpp:173-293, :536-595) ----
    s.n_up.assign(L, 0.0);
    s.n_dn.assign(L, 0.0);
    const Value& ic = sys["IC"];
    if (!ic.isNull()) {
        std::string type = ic.get("type", Value("givenDiagonal")).asString();
        s.ic_type = type;
        int64_t nUp = ic.get("nPartUp", Value((int64_t)std::ceil(L / 2.0))).asInt();
        int64_t nDn = ic.get("nPartDown", Value((int64_t)std::floor(L / 2.0))).asInt();
        if (type == "givenDiagonal") {
            std::string sub = ic.get("subtype", Value("spinWave")).asString();
            s.ic_type += "/" + sub;
            // particle numbers: nPart for spinWave / chargeWave (:536-545), the occupied counts for
            // spinWave3 / chargeWave3 (:196-227), none set for givenInJson / givenInFile
            if (sub == "spinWave" || sub == "chargeWave") {
                s.part_up = (int)std::min<int64_t>(nUp, 1 << 30);
                s.part_dn = (int)std::min<int64_t>(nDn, 1 << 30);
            }
            if (sub == "spinWave") {
                for (int64_t i = 0; i < std::min<int64_t>(L, 2 * nUp); i += 2) s.n_up[i] = 1.0;
                for (int64_t i = 1; i < std::min<int64_t>(L, 2 * nDn); i += 2) s.n_dn[i] = 1.0;
            } else if (sub == "chargeWave") {
                for (int64_t i = 0; i < std::min<int64_t>(L, 2 * nUp); i += 2) s.n_up[i] = 1.0;
                for (int64_t i = 0; i < std::min<int64_t>(L, 2 * nDn); i += 2) s.n_dn[i] = 1.0;
            } else if (sub == "spinWave3" || sub == "chargeWave3") {
                int mid = L / 2;  // std::ceil(getnSites()/2) on an integer quotient
                s.part_up = mid;
                s.part_dn = sub == "spinWave3" ? L - mid : mid;
                for (int i = 0; i < mid; ++i) s.n_up[i] = 1.0;
                if (sub == "spinWave3")
                    for (int i = mid; i < L; ++i) s.n_dn[i] = 1.0;
                else
                    for (int i = 0; i < mid; ++i) s.n_dn[i] = 1.0;
            } else if (sub == "givenInJson") {
                const Value& nu = ic["IC_NuDiag"];
                const Value& nd = ic["IC_NdDiag"];
                if (nu.isArray())
                    for (int i = 0; i < L && i < (int)nu.size(); ++i) s.n_up[i] = nu[i].asBool() ? 1.0 : 0.0;
                if (nd.isArray())
                    for (int i = 0; i < L && i < (int)nd.size(); ++i) s.n_dn[i] = nd[i].asBool() ? 1.0 : 0.0;
            } else if (sub == "givenInFile") {
                s.warnings.push_back("IC subtype givenInFile reads no file (as in the reference): empty state");
            }
        } else if (type == "tbGroundState" || type == "mfGroundState") {
            // setInitialConditions(type, nPart) (src/FHSystem.cpp:536-595) keeps nPartUp/nPartDown and
            // computes the eigenbasis; initiateSolver (src/FHfTWASolver.cpp:120-153) occupies the lowest
            // nPart states and rotates the sampled state with the eigenvector matrix.
            if (nUp < 0 || nDn < 0 || nUp > L || nDn > L) {
                if (err) *err = "initial condition '" + type + "': nPartUp/nPartDown must lie in [0, " + std::to_string(L) + "]";
                return false;
            }
            s.part_up = (int)nUp;
            s.part_dn = (int)nDn;
            if (!computeGroundState(s, type == "mfGroundState", (int)nUp, (int)nDn, err)) return false;
            const std::vector<double>& ru = type == "mfGroundState" ? s.mf_rot_up : s.tb_rot;
            const std::vector<double>& rd = type == "mfGroundState" ? s.mf_rot_dn : s.tb_rot;
            for (int64_t i = 0; i < nUp; ++i) s.n_up[i] = 1.0;
            for (int64_t i = 0; i < nDn; ++i) s.n_dn[i] = 1.0;
            s.rot_up.assign(2 * (size_t)L * L, 0.0);
            s.rot_dn.assign(2 * (size_t)L * L, 0.0);
            for (size_t k = 0; k < (size_t)L * L; ++k) {
                s.rot_up[2 * k] = ru[k];
                s.rot_dn[2 * k] = rd[k];
            }
        } else if (type == "groundState") {
            // the reference computes the mean-field state for this type but initiateSolver has no
            // branch for it (src/FHfTWASolver.cpp:99-153), leaving the rotation matrices empty
            if (err) *err = "initial condition 'groundState' has no fTWA initial state (use tbGroundState or mfGroundState)";
            return false;
        } else {
            s.warnings.push_back("initial condition type '" + type + "' sets no occupation");
        }
    } else {
        s.warnings.push_back("No Initial condition in the json file.");
    }
    if (sys.isMember("coreHole") && sys["coreHole"].get("presence", Value(false)).asBool()) {
        if (!addCoreHole(sys["coreHole"], s, err)) return false;
    }
    out = std::move(s);
    return true;
}

}  // namespace fpsamd
