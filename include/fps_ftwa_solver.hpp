// fps_ftwa_solver.hpp -- C++ facade with the reference's FHSystem / FHfTWASolver interface.
//
// Source-compatible with the way fps.cpp:192-207 drives the reference solver:
//     FHfTWASolver s(&fhSystem);
//     s.setResultFile(simName + "_fTWA_res", workFolder);
//     s.setJsonFile(rootSim);  s.setResultsWriter("json");
//     s.setResultsToSave(rootSim["results"]);  s.setMethodParameters(rootSim["parameters"]);
//     s.initiateSolver();  s.runSimulation();
// Method names and argument meaning follow include/FHfTWASolver.h:28-41 and include/FHSolver.h:23-47;
// the JSON value type is fpsjson::Value instead of Json::Value.  Everything below runs on the C ABI
// of fps_ftwa.h, i.e. on the gfx950 device; there is no CPU path.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "fps_json.hpp"
#include "fps_system.hpp"
#include "fps_ftwa.h"

// FHSystem(Json, folder) -- src/FHSystem.cpp:123-326 (the parts the fTWA path reads).
class FHSystem {
public:
    FHSystem() = default;
    // Throws std::runtime_error when the system cannot be built.
    FHSystem(const fpsjson::Value& systemJson, const std::string& folderName);
    unsigned getnSites() const { return static_cast<unsigned>(sys_.L); }
    const fpsamd::FtwaSystem& data() const { return sys_; }

private:
    fpsamd::FtwaSystem sys_;
};

class FHfTWASolver {
public:
    explicit FHfTWASolver(FHSystem* fhSystem);
    ~FHfTWASolver();
    FHfTWASolver(const FHfTWASolver&) = delete;
    FHfTWASolver& operator=(const FHfTWASolver&) = delete;
    FHfTWASolver(FHfTWASolver&&) noexcept;

    void setIntegratorOrder(int order);                 // accepted; RK4 "3/8" is always used (src/FHfTWASolver.cpp:90)
    void setfTWAParameters(unsigned int nTraj);         // src/FHfTWASolver.cpp:66-69
    void setMethodParameters(const fpsjson::Value& p);  // src/FHfTWASolver.cpp:49-60 (+ optional "mi355x" object)
    void setTimeParameters(double t_max, double dt, unsigned nPerTS = 1);  // src/FHSolver.cpp:95-101
    void setResultsToSave(const fpsjson::Value& res);   // src/FHSolver.cpp:34-45
    void setResultFile(std::string workName, std::string folderName);      // src/FHSolver.cpp:11-16
    void setResultsWriter(std::string how, bool yn = true);                // src/FHSolver.cpp:18-32
    void setJsonFile(const fpsjson::Value& file);       // src/FHSolver.cpp:109-112
    void setTalkative(bool yn) { talkative_ = yn; }

    void initiateSolver();  // src/FHfTWASolver.cpp:74-213
    void runSimulation();   // src/FHfTWASolver.cpp:222-556 (writes the results file when the json writer is on)

    // Engine controls with no reference counterpart.
    void setSeed(uint64_t seed) { seed_ = seed; }
    void setEngine(int engine) { engine_ = engine; }
    // The "simulation" object of the results file after runSimulation().
    const fpsjson::Value& results() const { return sim_; }
    // Whole results document (input echo + "fpsVersion" + "tbGroundState"/"mfGroundState" when an
    // eigenbasis was computed + "simulation" + "mi355x").
    fpsjson::Value resultsDocument() const;
    fpsftwa_handle* handle() const { return h_; }

protected:
    void check(int rc, const char* what) const;
    void writeResults();

    FHSystem* sys_ = nullptr;
    fpsftwa_handle* h_ = nullptr;
    fpsjson::Value json_ = fpsjson::Value::object();
    fpsjson::Value sim_;
    fpsjson::Value info_;
    fpsjson::Value gsTB_, gsMF_;           // "tbGroundState" / "mfGroundState" blocks, null when not computed
    std::string workName_, folderName_;
    bool writeJson_ = false, talkative_ = false, initiated_ = false;
    double tMax_ = 1.0, dt_ = 0.01;        // include/FHSolver.h:179-180 defaults
    unsigned nSub_ = 1;
    unsigned nTraj_ = 1;                   // include/FHfTWASolver.h:52
    uint32_t flags_ = FPSFTWA_RES_DIAG | FPSFTWA_RES_DIAG2 | FPSFTWA_RES_ENERGIES;  // include/FHSolver.h:158-164
    uint64_t seed_ = 0;
    int engine_ = FPSFTWA_ENGINE_FP64_BATCHED;
    bool saveTB_ = false, saveMF_ = false;  // "tbStates" / "mfStates" (src/FHSolver.cpp:41-42)
    int basisTB_ = -1, basisMF_ = -1;       // their index among the engine's bases
    int method_ = FPSFTWA_METHOD_FTWA;
};

// FHHFSolver (include/FHHFSolver.h, src/FHHFSolver.cpp): time-dependent Hartree-Fock, one
// deterministic density matrix with the Fock exchange of the V interaction, on the plain FP64
// engine (one state, so none of the batched engines' parallelism applies).
class FHHFSolver : public FHfTWASolver {
public:
    explicit FHHFSolver(FHSystem* fhSystem);
    void setMethodParameters(const fpsjson::Value& p);  // src/FHHFSolver.cpp:47-57 (dt, t_max, stepsParTimePoints, int_order)
};

// FHGPSRSolver (include/FHGPSRSolver.h, src/FHGPSRSolver.cpp): stochastic trajectories of the normally ordered
// density matrix (Gaussian phase-space representation).  Same interface and method-parameter defaults as
// FHfTWASolver (src/FHGPSRSolver.cpp:48-58); the batched FP64 engine by default (L <= 32, short-range hopping,
// no pulse), the plain FP64 engine otherwise.  Results file "<sim>_GPSR_res.json" (fps.cpp:213).
class FHGPSRSolver : public FHfTWASolver {
public:
    explicit FHGPSRSolver(FHSystem* fhSystem);
};
