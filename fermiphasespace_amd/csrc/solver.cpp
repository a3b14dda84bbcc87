// solver.cpp -- FHSystem / FHfTWASolver facade, the results-file writer, and the JSON-level
// C ABI entry points (fpsftwa_run_json_file, fpsftwa_load_system_json).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "../../include/fps_ftwa_solver.hpp"
#include "eigen.hpp"
#include "ftwa_internal.hpp"

using fpsjson::Value;

namespace {

// Version of the reference format the results file follows (include/FHSystem.h:21-23).
const char* kFormatVersion = "0.5.0";

Value rowOf(const std::vector<double>& a, size_t row, size_t T)
{
    Value v = Value::array();
    for (size_t t = 0; t < T; ++t) v.append(a[row * T + t]);
    return v;
}

Value rowsOf(const std::vector<double>& a, size_t rows, size_t T)
{
    Value v = Value::array();
    for (size_t r = 0; r < rows; ++r) v.append(rowOf(a, r, T));
    return v;
}

std::vector<double> fetch(fpsftwa_handle* h, int which)
{
    size_t n = 0;
    if (fpsftwa_get_result(h, which, nullptr, 0, &n) != FPSFTWA_OK) return {};
    std::vector<double> v(n);
    if (fpsftwa_get_result(h, which, v.data(), n, nullptr) != FPSFTWA_OK) return {};
    return v;
}

const char* engineName(int e)
{
    switch (e) {
    case FPSFTWA_ENGINE_FP64_PLAIN: return "fp64_plain";
    case FPSFTWA_ENGINE_FP64_BATCHED: return "fp64";
    default: return "fp32";
    }
}

}  // namespace

FHHFSolver::FHHFSolver(FHSystem* fhSystem) : FHfTWASolver(fhSystem)
{
    method_ = FPSFTWA_METHOD_HF;
    nTraj_ = 1;
    engine_ = FPSFTWA_ENGINE_FP64_PLAIN;
}

FHGPSRSolver::FHGPSRSolver(FHSystem* fhSystem) : FHfTWASolver(fhSystem)
{
    method_ = FPSFTWA_METHOD_GPSR;
}

void FHHFSolver::setMethodParameters(const Value& p)
{
    // defaults of src/FHHFSolver.cpp:47-57; RK4 "3/8" is forced in initiateSolver (:199)
    setTimeParameters(p.get("t_max", Value(2.0)).asDouble(), p.get("dt", Value(0.01)).asDouble(),
                      static_cast<unsigned>(std::max<int64_t>(1, p.get("stepsParTimePoints", Value(1)).asInt())));
    setIntegratorOrder(static_cast<int>(p.get("int_order", Value(2)).asInt()));
}

FHSystem::FHSystem(const Value& systemJson, const std::string& folderName)
{
    std::string err;
    if (!fpsamd::buildSystem(systemJson, folderName, sys_, &err)) throw std::runtime_error(err);
}

FHfTWASolver::FHfTWASolver(FHSystem* fhSystem) : sys_(fhSystem)
{
    if (!sys_) throw std::invalid_argument("FHfTWASolver needs a system");
}

FHfTWASolver::FHfTWASolver(FHfTWASolver&& o) noexcept
    : sys_(o.sys_), h_(o.h_), json_(std::move(o.json_)), sim_(std::move(o.sim_)), info_(std::move(o.info_)),
      gsTB_(std::move(o.gsTB_)), gsMF_(std::move(o.gsMF_)), workName_(std::move(o.workName_)), folderName_(std::move(o.folderName_)), writeJson_(o.writeJson_),
      talkative_(o.talkative_), initiated_(o.initiated_), tMax_(o.tMax_), dt_(o.dt_), nSub_(o.nSub_), nTraj_(o.nTraj_),
      flags_(o.flags_), seed_(o.seed_), engine_(o.engine_), saveTB_(o.saveTB_), saveMF_(o.saveMF_),
      basisTB_(o.basisTB_), basisMF_(o.basisMF_), method_(o.method_)
{
    o.h_ = nullptr;
}

FHfTWASolver::~FHfTWASolver()
{
    if (h_) fpsftwa_destroy(h_);
}

void FHfTWASolver::check(int rc, const char* what) const
{
    if (rc != FPSFTWA_OK) throw std::runtime_error(std::string(what) + ": " + fpsftwa_last_error());
}

void FHfTWASolver::setIntegratorOrder(int order)
{
    if (order != 2 && order != 4)
        std::cout << "Integrator order " << order << " not implemented. Switching to default 2." << std::endl;
}

void FHfTWASolver::setfTWAParameters(unsigned int nTraj) { nTraj_ = nTraj; }

void FHfTWASolver::setTimeParameters(double t_max, double dt, unsigned nPerTS)
{
    tMax_ = t_max;
    dt_ = dt;
    nSub_ = nPerTS;
}

void FHfTWASolver::setMethodParameters(const Value& p)
{
    const double dt = p.get("dt", Value(0.01)).asDouble();
    const double tm = p.get("t_max", Value(2.0)).asDouble();
    // the reference reads "stepsParTimePoints" (src/FHfTWASolver.cpp:53); the documented
    // "stepsPerTimePoints" is ignored there, and here too
    const int64_t sptp = p.get("stepsParTimePoints", Value(1)).asInt();
    if (p.isMember("stepsPerTimePoints") && !p.isMember("stepsParTimePoints"))
        std::cout << "   -- note: 'stepsPerTimePoints' is not read by the fTWA solver; 'stepsParTimePoints' is."
                  << std::endl;
    const int64_t nTraj = p.get("n_traj", Value(10)).asInt();
    setTimeParameters(tm, dt, static_cast<unsigned>(std::max<int64_t>(1, sptp)));
    setIntegratorOrder(static_cast<int>(p.get("int_order", Value(2)).asInt()));
    setfTWAParameters(static_cast<unsigned>(std::max<int64_t>(1, nTraj)));
    const Value& m = p["mi355x"];
    if (m.isObject()) {
        seed_ = static_cast<uint64_t>(m.get("seed", Value(0)).asInt());
        const std::string e = m.get("engine", Value("fp64")).asString();
        if (e == "fp32") engine_ = FPSFTWA_ENGINE_FP32_BATCHED;
        else if (e == "fp64_plain") engine_ = FPSFTWA_ENGINE_FP64_PLAIN;
        else engine_ = FPSFTWA_ENGINE_FP64_BATCHED;
    }
}

void FHfTWASolver::setResultsToSave(const Value& res)
{
    uint32_t f = 0;
    if (res.get("diagDMatrix", Value(true)).asBool()) f |= FPSFTWA_RES_DIAG;
    if (res.get("diag2DMatrix", Value(false)).asBool()) f |= FPSFTWA_RES_DIAG2;
    if (res.get("onSiteDoublons", Value(false)).asBool()) f |= FPSFTWA_RES_DOUBLONS;
    if (res.get("wholeDMatrix", Value(false)).asBool()) f |= FPSFTWA_RES_WHOLE;
    if (res.get("energies", Value(true)).asBool()) f |= FPSFTWA_RES_ENERGIES;
    saveTB_ = res.get("tbStates", Value(false)).asBool();
    saveMF_ = res.get("mfStates", Value(false)).asBool();
    flags_ = f;
}

void FHfTWASolver::setResultFile(std::string workName, std::string folderName)
{
    workName_ = std::move(workName);
    folderName_ = std::move(folderName);
    talkative_ = false;
}

void FHfTWASolver::setResultsWriter(std::string how, bool yn)
{
    if (how == "json") writeJson_ = yn;
    else if (how == "txt" || how == "hdf5")
        std::cout << "   -- note: the '" << how << "' writer is not provided; use json." << std::endl;
}

void FHfTWASolver::setJsonFile(const Value& file) { json_ = file; }

void FHfTWASolver::initiateSolver()
{
    std::cout << (method_ == FPSFTWA_METHOD_GPSR ? "   -- GPSR Solver initiation." : "   -- fTWA Solver initiation.")
              << std::endl;
    const fpsamd::FtwaSystem& s = sys_->data();
    if (!h_) {
        fpsftwa_system d{};
        d.n_sites = s.L;
        d.hopping = s.tM.data();
        d.coulomb_type = s.coulomb_type;
        d.coulomb_u = s.U.data();
        d.coulomb_v = s.V.data();
        d.potential = s.potential.empty() ? nullptr : s.potential.data();
        d.pulse_type = s.pulse_type;
        d.pulse_A0 = s.pulse_A0;
        d.pulse_t0 = s.pulse_t0;
        d.pulse_td = s.pulse_td;
        d.pulse_w0 = s.pulse_w0;
        d.pulse_phi = s.pulse_phi;
        for (int k = 0; k < 3; ++k) d.pulse_pol[k] = s.pulse_pol[k];
        d.pulse_n_knots = static_cast<int32_t>(s.pulse_knots.size());
        d.pulse_knots = s.pulse_knots.empty() ? nullptr : s.pulse_knots.data();
        d.displacement = s.displacement.data();
        check(fpsftwa_create(&d, &h_), "fpsftwa_create");
    }
    check(fpsftwa_set_method(h_, method_), "method");
    check(fpsftwa_set_initial_state(h_, s.n_up.data(), s.n_dn.data(), s.rot_up.empty() ? nullptr : s.rot_up.data(),
                                    s.rot_dn.empty() ? nullptr : s.rot_dn.data()),
          "initial state");
    // bases of the "tbStates" / "mfStates" observables: the eigenbases the system already holds
    // (eigenbasis initial conditions), else computed here as initiateSolver does
    // (src/FHfTWASolver.cpp:80-87) for the system's particle numbers
    const size_t LL = (size_t)s.L * s.L;
    std::vector<double> bases;
    basisTB_ = basisMF_ = -1;
    // eigenbases the system computed for its initial state go into the results file as
    // "tbGroundState" / "mfGroundState" (src/FHSolver.cpp:177-185), as do those computed below
    gsTB_ = s.tb_rot.empty() ? Value() : fpsamd::tbGroundStateJson(s.tb_energy, s.tb_rot);
    gsMF_ = s.mf_rot_up.empty() ? Value()
                                : fpsamd::mfGroundStateJson(s.mf_energy_up, s.mf_energy_dn, s.mf_rot_up, s.mf_rot_dn,
                                                    s.mf_iterations, s.mf_residue);
    auto add_basis = [&](bool mean_field) {
        std::vector<double> ru, rd;
        if (!mean_field && !s.tb_rot.empty()) {
            ru = rd = s.tb_rot;
        } else if (mean_field && !s.mf_rot_up.empty()) {
            ru = s.mf_rot_up;
            rd = s.mf_rot_dn;
        } else {
            if (s.part_up < 0 || s.part_dn < 0 || s.part_up > s.L || s.part_dn > s.L)
                throw std::runtime_error("particle numbers outside [0, L] for the ground-state bases");
            ru.assign(LL, 0.0);
            rd.assign(LL, 0.0);
            std::vector<double> eu(s.L), ed(s.L);
            int it = 0;
            double res = 0.0;
            if (fpsamd::ground_state_rotation(s.L, s.tM.data(), s.U.data(), s.part_up, s.part_dn, mean_field, ru.data(),
                                              rd.data(), &it, &res, eu.data(), ed.data()) != 0)
                throw std::runtime_error("eigen-decomposition of the hopping matrix did not converge");
            if (mean_field && it >= 100)
                std::cout << "!! the MF ground state computation has not converged. Residue=" << res << std::endl;
            if (mean_field) gsMF_ = fpsamd::mfGroundStateJson(eu, ed, ru, rd, it, res);
            else gsTB_ = fpsamd::tbGroundStateJson(eu, ru);
        }
        bases.insert(bases.end(), ru.begin(), ru.end());
        bases.insert(bases.end(), rd.begin(), rd.end());
    };
    if (s.core_hole_site >= 0) {
        // NiCh = <rho_sigma(site, L-1)> (src/FHfTWASolver.cpp:413-417); it takes the engine's recorded
        // element, which wholeDMatrix's corner (0, L-1) also needs unless the hole is linked to site 0
        if ((flags_ & FPSFTWA_RES_WHOLE) && s.core_hole_site != 0)
            throw std::runtime_error("wholeDMatrix together with a core hole linked to site " +
                                     std::to_string(s.core_hole_site) + " is not supported by this build");
        if (saveTB_ || saveMF_)  // the reference's L x L bases cannot act on the L+1 hole system either
            throw std::runtime_error("tbStates / mfStates are not defined for a core-hole system");
        check(fpsftwa_set_recorded_element(h_, s.core_hole_site, s.L - 1), "core-hole element");
    }
    int nb = 0;
    if (saveTB_) {
        add_basis(false);
        basisTB_ = nb++;
    }
    if (saveMF_) {
        add_basis(true);
        basisMF_ = nb++;
    }
    check(fpsftwa_set_bases(h_, nb, nb ? bases.data() : nullptr), "bases");
    initiated_ = true;
}

void FHfTWASolver::runSimulation()
{
    if (!initiated_) initiateSolver();
    std::cout << "   -- Start simulation." << std::endl;
    check(fpsftwa_set_params(h_, dt_, tMax_, static_cast<int32_t>(nSub_), nTraj_, 4, seed_, engine_), "parameters");
    check(fpsftwa_set_results(h_, flags_), "results");
    int rc = fpsftwa_run(h_);
    if (rc == FPSFTWA_EUNSUPPORTED && method_ == FPSFTWA_METHOD_GPSR && engine_ != FPSFTWA_ENGINE_FP64_PLAIN) {
        // the batched GPSR kernel covers small lattices (L <= 32, short-range hopping, no pulse); the plain
        // FP64 engine runs the rest
        std::cout << "   -- note: " << fpsftwa_last_error() << "; running GPSR on the fp64_plain engine." << std::endl;
        engine_ = FPSFTWA_ENGINE_FP64_PLAIN;
        check(fpsftwa_set_params(h_, dt_, tMax_, static_cast<int32_t>(nSub_), nTraj_, 4, seed_, engine_), "parameters");
        rc = fpsftwa_run(h_);
    }
    check(rc, "run");
    const int L = sys_->data().L;
    std::vector<double> t = fetch(h_, FPSFTWA_OUT_TIME);
    const size_t T = t.size();
    Value sim = Value::object();
    Value tv = Value::array();
    for (double x : t) tv.append(x);
    sim["t"] = tv;
    if (flags_ & FPSFTWA_RES_WHOLE) {  // the single corner element N(0, L-1)(t) (src/FHSolver.cpp:196-197)
        for (int which : {FPSFTWA_OUT_N1HU, FPSFTWA_OUT_N1HD}) {
            std::vector<double> c = fetch(h_, which);
            Value o = Value::object();
            o["r"] = rowOf(c, 0, T);
            o["i"] = rowOf(c, 1, T);
            sim[which == FPSFTWA_OUT_N1HU ? "N1hu" : "N1hd"] = o;
        }
    }
    if (flags_ & FPSFTWA_RES_DIAG) {
        sim["Niiu"] = rowsOf(fetch(h_, FPSFTWA_OUT_NIIU), L, T);
        sim["Niid"] = rowsOf(fetch(h_, FPSFTWA_OUT_NIID), L, T);
    }
    if (flags_ & FPSFTWA_RES_DIAG2) {
        sim["N2iiu"] = rowsOf(fetch(h_, FPSFTWA_OUT_N2IIU), L, T);
        sim["N2iid"] = rowsOf(fetch(h_, FPSFTWA_OUT_N2IID), L, T);
    }
    if (flags_ & FPSFTWA_RES_DOUBLONS) sim["NuNd"] = rowsOf(fetch(h_, FPSFTWA_OUT_NUND), L, T);
    if (sys_->data().core_hole_site >= 0) {  // "NiCh": {"u": {r, i}, "d": {r, i}} (src/FHSolver.cpp:285-291)
        Value nich = Value::object();
        for (int which : {FPSFTWA_OUT_N1HU, FPSFTWA_OUT_N1HD}) {
            std::vector<double> c = fetch(h_, which);
            Value o = Value::object();
            o["r"] = rowOf(c, 0, T);
            o["i"] = rowOf(c, 1, T);
            nich[which == FPSFTWA_OUT_N1HU ? "u" : "d"] = o;
        }
        sim["NiCh"] = nich;
    }
    if (basisTB_ >= 0 || basisMF_ >= 0) {  // src/FHSolver.cpp:242-265
        const std::vector<double> m = fetch(h_, FPSFTWA_OUT_BASIS);
        auto block = [&](int b, int sg) {
            std::vector<double> v(m.begin() + ((size_t)(b * 2 + sg) * L) * T,
                                  m.begin() + ((size_t)(b * 2 + sg + 1) * L) * T);
            return rowsOf(v, L, T);
        };
        if (basisTB_ >= 0) {
            sim["Mkku"] = block(basisTB_, 0);
            sim["Mkkd"] = block(basisTB_, 1);
        }
        if (basisMF_ >= 0) {
            sim["Mukku"] = block(basisMF_, 0);
            sim["Mukkd"] = block(basisMF_, 1);
        }
    }
    if (flags_ & FPSFTWA_RES_ENERGIES) {
        std::vector<double> e = fetch(h_, FPSFTWA_OUT_ENERGY);
        Value E = Value::object();
        E["tH"] = rowOf(e, 0, T);
        E["U"] = rowOf(e, 1, T);
        E["V"] = rowOf(e, 2, T);
        E["pe"] = rowOf(e, 3, T);
        E["tot"] = rowOf(e, 4, T);
        sim["E"] = E;
    }
    sim_ = sim;
    int64_t launches = 0;
    double kms = 0, rms = 0;
    fpsftwa_get_timing(h_, &launches, &kms, &rms);
    info_ = Value::object();
    info_["engine"] = engineName(engine_);
    if (method_ == FPSFTWA_METHOD_HF) info_["method"] = "HF (one deterministic state, Fock exchange with V)";
    if (method_ == FPSFTWA_METHOD_GPSR)
        info_["method"] = "GPSR (normally ordered density matrix, analytic diffusion of variance h per step)";
    info_["seed"] = static_cast<int64_t>(seed_);
    info_["kernel_ms"] = kms;
    info_["run_ms"] = rms;
    info_["traj_steps_per_s"] =
        rms > 0 ? static_cast<double>(nTraj_) * static_cast<double>(T) * nSub_ / (rms * 1e-3) : 0.0;
    info_["energies_note"] = "E is computed for every system (the reference writes E only for core-hole systems)";
    Value err = Value::object();
    if (flags_ & FPSFTWA_RES_DIAG) {
        err["Niiu"] = rowsOf(fetch(h_, FPSFTWA_OUT_NIIU_ERR), L, T);
        err["Niid"] = rowsOf(fetch(h_, FPSFTWA_OUT_NIID_ERR), L, T);
    }
    if (flags_ & FPSFTWA_RES_DOUBLONS) err["NuNd"] = rowsOf(fetch(h_, FPSFTWA_OUT_NUND_ERR), L, T);
    info_["statistical_error"] = err;
    if (writeJson_) writeResults();
}

Value FHfTWASolver::resultsDocument() const
{
    Value res = json_;
    if (!res.isObject()) res = Value::object();
    res["fpsVersion"] = kFormatVersion;
    if (!gsTB_.isNull()) res["tbGroundState"] = gsTB_;  // src/FHSolver.cpp:177-185
    if (!gsMF_.isNull()) res["mfGroundState"] = gsMF_;
    res["simulation"] = sim_;
    res["mi355x"] = info_;
    return res;
}

void FHfTWASolver::writeResults()
{
    const std::string path = folderName_ + workName_ + ".json";
    std::ofstream f(path);
    if (!f.good()) throw std::runtime_error("cannot write " + path);
    f << fpsjson::write(resultsDocument(), "\t");
    std::cout << "   -- Results written in " << folderName_ << workName_ << ".json" << std::endl;
}

// ---- JSON-level C ABI ---------------------------------------------------------------------------
namespace {

bool splitPath(const std::string& path, std::string& folder, std::string& file)
{
    const size_t p = path.find_last_of('/');
    folder = (p == std::string::npos) ? "" : path.substr(0, p + 1);
    file = (p == std::string::npos) ? path : path.substr(p + 1);
    return true;
}

}  // namespace

extern "C" int fpsftwa_run_json_file(const char* path, const char* out_path)
{
    if (!path) {
        fpsamd::set_last_error("null path");
        return FPSFTWA_EINVAL;
    }
    try {
        Value root;
        std::string err;
        if (!fpsjson::parseFile(path, root, &err)) {
            fpsamd::set_last_error(err);
            return FPSFTWA_EIO;
        }
        std::string folder, file;
        splitPath(path, folder, file);
        const std::string simName = root.get("sim_name", Value("defaultSimName")).asString();
        const std::string method = root.get("method", Value("HF")).asString();
        if (method != "fTWA" && method != "HF" && method != "GPSR") {
            fpsamd::set_last_error("method '" + method + "' is not provided by this engine (fTWA, HF and GPSR are)");
            return FPSFTWA_EUNSUPPORTED;
        }
        FHSystem sys(root["system"], folder);
        for (const auto& w : sys.data().warnings) std::cout << "   -- note: " << w << std::endl;
        // fps.cpp:88-220: the HF, fTWA and GPSR branches differ only in the solver class and the file suffix
        std::unique_ptr<FHfTWASolver> owner(method == "HF"     ? static_cast<FHfTWASolver*>(new FHHFSolver(&sys))
                                            : method == "GPSR" ? static_cast<FHfTWASolver*>(new FHGPSRSolver(&sys))
                                                               : new FHfTWASolver(&sys));
        FHfTWASolver& solver = *owner;
        std::string work = simName + "_" + method + "_res";
        std::string outFolder = folder;
        if (out_path) {
            std::string f2, w2;
            splitPath(out_path, f2, w2);
            outFolder = f2;
            work = w2.size() > 5 && w2.compare(w2.size() - 5, 5, ".json") == 0 ? w2.substr(0, w2.size() - 5) : w2;
        }
        solver.setResultFile(work, outFolder);
        solver.setJsonFile(root);
        solver.setResultsWriter("json");
        solver.setResultsToSave(root["results"]);
        if (method == "HF") static_cast<FHHFSolver&>(solver).setMethodParameters(root["parameters"]);
        else solver.setMethodParameters(root["parameters"]);
        solver.initiateSolver();
        solver.runSimulation();
    } catch (const std::exception& e) {
        fpsamd::set_last_error(e.what());
        return FPSFTWA_EINVAL;
    }
    return FPSFTWA_OK;
}

extern "C" int fpsftwa_load_system_json(const char* path, int32_t* n_sites, int32_t* coulomb_type, double* hopping,
                                        double* coulomb_u, double* coulomb_v, double* potential, double* n_up,
                                        double* n_dn, double* positions)
{
    if (!path || !n_sites) {
        fpsamd::set_last_error("null argument");
        return FPSFTWA_EINVAL;
    }
    Value root;
    std::string err;
    if (!fpsjson::parseFile(path, root, &err)) {
        fpsamd::set_last_error(err);
        return FPSFTWA_EIO;
    }
    std::string folder, file;
    splitPath(path, folder, file);
    fpsamd::FtwaSystem s;
    if (!fpsamd::buildSystem(root["system"], folder, s, &err)) {
        fpsamd::set_last_error(err);
        return FPSFTWA_EINVAL;
    }
    *n_sites = s.L;
    if (coulomb_type) *coulomb_type = s.coulomb_type;
    const size_t LL = (size_t)s.L * s.L;
    if (hopping) std::memcpy(hopping, s.tM.data(), LL * sizeof(double));
    if (coulomb_u) std::memcpy(coulomb_u, s.U.data(), LL * sizeof(double));
    if (coulomb_v) std::memcpy(coulomb_v, s.V.data(), LL * sizeof(double));
    if (potential) {
        for (int i = 0; i < s.L; ++i) potential[i] = s.potential.empty() ? 0.0 : s.potential[i];
    }
    if (n_up) std::memcpy(n_up, s.n_up.data(), s.L * sizeof(double));
    if (n_dn) std::memcpy(n_dn, s.n_dn.data(), s.L * sizeof(double));
    if (positions) std::memcpy(positions, s.positions.data(), 3 * (size_t)s.L * sizeof(double));
    return FPSFTWA_OK;
}

extern "C" int fpsftwa_load_initial_state_json(const char* path, int32_t* n_sites, double* n_up, double* n_dn,
                                               double* rot_up, double* rot_dn)
{
    if (!path || !n_sites) {
        fpsamd::set_last_error("null argument");
        return FPSFTWA_EINVAL;
    }
    Value root;
    std::string err;
    if (!fpsjson::parseFile(path, root, &err)) {
        fpsamd::set_last_error(err);
        return FPSFTWA_EIO;
    }
    std::string folder, file;
    splitPath(path, folder, file);
    fpsamd::FtwaSystem s;
    if (!fpsamd::buildSystem(root["system"], folder, s, &err)) {
        fpsamd::set_last_error(err);
        return FPSFTWA_EINVAL;
    }
    *n_sites = s.L;
    const size_t LL = (size_t)s.L * s.L;
    if (n_up) std::memcpy(n_up, s.n_up.data(), s.L * sizeof(double));
    if (n_dn) std::memcpy(n_dn, s.n_dn.data(), s.L * sizeof(double));
    for (int spin = 0; spin < 2; ++spin) {
        double* out = spin == 0 ? rot_up : rot_dn;
        const std::vector<double>& r = spin == 0 ? s.rot_up : s.rot_dn;
        if (!out) continue;
        if (!r.empty()) {
            std::memcpy(out, r.data(), 2 * LL * sizeof(double));
        } else {
            std::memset(out, 0, 2 * LL * sizeof(double));
            for (int i = 0; i < s.L; ++i) out[2 * ((size_t)i * s.L + i)] = 1.0;
        }
    }
    return FPSFTWA_OK;
}

extern "C" int fpsftwa_ground_state(int32_t n_sites, const double* hopping, const double* coulomb_u, int32_t n_up,
                                    int32_t n_dn, int32_t mean_field, double* rot_up, double* rot_dn,
                                    int32_t* iterations, double* residue, double* energy_up, double* energy_dn)
{
    if (n_sites <= 0 || !hopping || !rot_up || !rot_dn || (mean_field && !coulomb_u)) {
        fpsamd::set_last_error("fpsftwa_ground_state: null argument or n_sites <= 0");
        return FPSFTWA_EINVAL;
    }
    if (n_up < 0 || n_dn < 0 || n_up > n_sites || n_dn > n_sites) {
        fpsamd::set_last_error("fpsftwa_ground_state: particle numbers must lie in [0, n_sites]");
        return FPSFTWA_EINVAL;
    }
    int it = 0;
    double res = 0.0;
    if (fpsamd::ground_state_rotation(n_sites, hopping, coulomb_u, n_up, n_dn, mean_field != 0, rot_up, rot_dn, &it,
                                      &res, energy_up, energy_dn) != 0) {
        fpsamd::set_last_error("fpsftwa_ground_state: eigen-decomposition did not converge");
        return FPSFTWA_EINVAL;
    }
    if (iterations) *iterations = it;
    if (residue) *residue = res;
    return FPSFTWA_OK;
}
