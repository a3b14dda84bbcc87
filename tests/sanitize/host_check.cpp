// host_check.cpp -- CPU driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of the host
// code that parses user input: the JSON reader and writer (csrc/json.cpp), the system builder
// (csrc/system.cpp) and the eigensolver (csrc/eigen.cpp).  Test infrastructure: built and run by
// tests/test_host_sanitize.py with g++ -fsanitize=address,undefined; no GPU code is involved.
//
//   host_check <simulation.json>...
// For every file that parses (malformed ones must fail cleanly): write, re-parse (the two values must be equal); build the system from its
// "system" object (positions file, interactions, initial state, core hole); compute the tight-binding and
// mean-field eigenbases for its particle numbers; then parse every prefix of the file's text at 64 cut
// points, each of which must either fail cleanly or give a value.  Exit status 0 when every step behaved.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../fermiphasespace_amd/csrc/eigen.hpp"
#include "../../include/fps_json.hpp"
#include "../../include/fps_system.hpp"

namespace {

int fails = 0;

void expect(bool ok, const std::string& what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        ++fails;
    }
}

std::string folderOf(const std::string& path)
{
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

}  // namespace

int main(int argc, char** argv)
{
    int parsed = 0, systems = 0, prefixes = 0;
    for (int a = 1; a < argc; ++a) {
        const std::string path = argv[a];
        std::ifstream f(path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        expect(!text.empty(), path + ": readable");
        fpsjson::Value v, w;
        std::string err;
        if (!fpsjson::parse(text, v, &err)) {  // malformed input: a clean failure is the expected outcome
            std::printf("note %s: does not parse: %s\n", path.c_str(), err.c_str());
            continue;
        }
        ++parsed;
        const std::string out = fpsjson::write(v, "\t");
        expect(fpsjson::parse(out, w, &err) && w == v, path + ": write / re-parse round trip");
        fpsamd::FtwaSystem s;
        if (fpsamd::buildSystem(v["system"], folderOf(path), s, &err)) {
            ++systems;
            expect(s.L > 0 && s.tM.size() == static_cast<size_t>(s.L) * s.L, path + ": system shape");
            if (s.core_hole_site < 0 && s.L <= 256) {
                for (int mf = 0; mf < 2; ++mf) {
                    const int nu = std::min(std::max(s.part_up, 0), s.L), nd = std::min(std::max(s.part_dn, 0), s.L);
                    fpsamd::FtwaSystem c = s;
                    expect(fpsamd::computeGroundState(c, mf == 1, nu, nd, &err),
                           path + ": ground state (" + err + ")");
                }
            }
        } else {
            std::printf("note %s: system not built: %s\n", path.c_str(), err.c_str());
        }
        const size_t n = text.size();
        for (int k = 0; k < 64; ++k) {  // truncated texts: a clean failure or a value, never a crash
            const size_t cut = n * static_cast<size_t>(k) / 64;
            fpsjson::Value p;
            std::string e2;
            (void)fpsjson::parse(text.substr(0, cut), p, &e2);
            ++prefixes;
        }
    }
    std::printf("host_check: %d files, %d parsed, %d systems, %d truncated parses, %d failures\n", argc - 1, parsed,
                systems, prefixes, fails);
    return fails == 0 ? 0 : 1;
}
