// lattice_gen.cpp -- the lattice generators and the system file of `makeSitesSystem` (fps_make_system).
//
// Restates makeSitesSystem.cpp:47-255 (arguments, defaults, the simulation file it writes), the
// generator constructor FHLattice(std::string, std::vector<double>) (src/FHLattice.cpp:169-321), the
// position generators and lattice boxes (src/FHLattice.cpp:860-1381), FHLattice::writePositions and
// writeFHLatticeJson (:715-771), FHSystem::setInitialConditions (src/FHSystem.cpp:453-533),
// setBasisInteractions (:599-603) and writeFHSystemJson (:1025-1140).  The site order of every generator
// follows the reference's loops, since it fixes the site numbering of the systems these files describe.
#include "../../include/fps_system.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

namespace fpsamd {

namespace {

using P3 = std::array<double, 3>;

const double kPi = 3.14159265358979323846;

// box matrices, row-major [3][3] (columns are the lattice vectors)
void diagBox(double nx, double ny, double nz, double a, double* m)
{
    std::fill(m, m + 9, 0.0);
    m[0] = nx * a;
    m[4] = ny * a;
    m[8] = nz * a;
}

// src/FHLattice.cpp:1071-1090: the sqrt(nCombs) x sqrt(nCombs) parallelogram of the comb sequence
void honeyCombBox(unsigned nCombs, double a, double* m)
{
    const double s3 = std::sqrt(3.0), r = std::sqrt(static_cast<double>(nCombs));
    const unsigned f = static_cast<unsigned>(std::floor(r)), c = static_cast<unsigned>(std::ceil(r));
    const unsigned nx = c, ny = f + ((nCombs - f * f) > f ? 1u : 0u);
    std::fill(m, m + 9, 0.0);
    m[0] = s3 * a * nx;
    m[1] = s3 / 2 * a * ny;
    m[4] = 1.5 * a * ny;
    m[8] = 10 * a;
}

void grapheneBox(int nx, int ny, double a, double* m)
{
    const double s3 = std::sqrt(3.0);
    std::fill(m, m + 9, 0.0);
    m[0] = s3 * a * nx;
    m[1] = s3 / 2 * a * ny;
    m[4] = 1.5 * a * ny;
    m[8] = 10 * a;
}

void tetragoneBox(int nx, int ny, int nz, double a, double* m)
{
    const double s3 = std::sqrt(3.0);
    std::fill(m, m + 9, 0.0);
    m[0] = a * nx;
    m[1] = a / 2 * ny;
    m[2] = a / 2 * nz;
    m[4] = s3 * a / 2 * ny;
    m[5] = a / (2 * s3) * nz;
    m[8] = std::sqrt(2.0) * a / s3 * nz;
}

std::vector<P3> cubicSites(unsigned nx, unsigned ny, unsigned nz, double a)
{
    std::vector<P3> p;
    for (unsigned z = 0; z < nz; ++z)
        for (unsigned y = 0; y < ny; ++y)
            for (unsigned x = 0; x < nx; ++x) p.push_back({a * x, a * y, a * z});
    return p;
}

// corners of the (nx+1)(ny+1)(nz+1) grid, each followed by the centre of the cube it opens
std::vector<P3> bccSites(unsigned nx, unsigned ny, unsigned nz, double a)
{
    std::vector<P3> p;
    for (unsigned z = 0; z <= nz; ++z)
        for (unsigned y = 0; y <= ny; ++y)
            for (unsigned x = 0; x <= nx; ++x) {
                p.push_back({a * x, a * y, a * z});
                if (x != nx && y != ny && z != nz) p.push_back({a * (x + 0.5), a * (y + 0.5), a * (z + 0.5)});
            }
    return p;
}

// corners, each followed by the centres of the three faces it opens (yz, xz, xy)
std::vector<P3> fccSites(unsigned nx, unsigned ny, unsigned nz, double a)
{
    std::vector<P3> p;
    for (unsigned z = 0; z <= nz; ++z)
        for (unsigned y = 0; y <= ny; ++y)
            for (unsigned x = 0; x <= nx; ++x) {
                p.push_back({a * x, a * y, a * z});
                if (y != ny && z != nz) p.push_back({a * x, a * (y + 0.5), a * (z + 0.5)});
                if (x != nx && z != nz) p.push_back({a * (x + 0.5), a * y, a * (z + 0.5)});
                if (x != nx && y != ny) p.push_back({a * (x + 0.5), a * (y + 0.5), a * z});
            }
    return p;
}

// the comb sequence (0,0), (1,0), (0,1), (1,1), (2,0), (2,1), (0,2), (1,2), (2,2), ... that keeps the
// first combs as square as possible (src/FHLattice.cpp:1335-1354)
void nextCombIndex(unsigned& i, unsigned& j)
{
    if (i == j) {
        i = i + 1;
        j = 0;
    } else if (i == j + 1) {
        j = i;
        i = 0;
    } else if (i > j) {
        j = j + 1;
    } else {
        i = i + 1;
    }
}

// the six corners of a comb centred at the origin, in the reference's order
std::array<P3, 6> combCorners(double a)
{
    const double h = std::sqrt(3.0) / 2;
    return {P3{-h * a, -0.5 * a, 0}, P3{-h * a, 0.5 * a, 0}, P3{0, a, 0},
            P3{0, -a, 0},            P3{h * a, -0.5 * a, 0}, P3{h * a, 0.5 * a, 0}};
}

// removes later copies of each point (at most two more: a honeycomb corner is shared by three combs),
// starting at point `from`, with the L1 distance below `threshold`
void removeCopies(std::vector<P3>& p, size_t from, double threshold)
{
    for (size_t i1 = from; i1 < p.size(); ++i1) {
        int found = 0;
        size_t i2 = i1 + 1;
        while (i2 < p.size() && found < 2) {
            const double d = std::fabs(p[i1][0] - p[i2][0]) + std::fabs(p[i1][1] - p[i2][1]) +
                             std::fabs(p[i1][2] - p[i2][2]);
            if (d < threshold) {
                p.erase(p.begin() + static_cast<std::ptrdiff_t>(i2));
                ++found;
            } else {
                ++i2;
            }
        }
    }
}

std::vector<P3> honeyCombSites(unsigned nCombs, double a)
{
    const double s3 = std::sqrt(3.0);
    const std::array<P3, 6> c0 = combCorners(a);
    std::vector<P3> p;
    unsigned bi = 0, bj = 0;
    for (unsigned c = 0; c < nCombs; ++c) {
        for (const P3& q : c0)
            p.push_back({q[0] + s3 * a * bi + s3 / 2 * a * bj, q[1] + 1.5 * a * bj, q[2]});
        nextCombIndex(bi, bj);
    }
    removeCopies(p, 0, 1e-4 * a);
    return p;
}

std::vector<P3> grapheneSites(int nx, int ny, double a)
{
    const double s3 = std::sqrt(3.0);
    std::vector<P3> p;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double x = s3 / 2 * a * j + s3 * a * i, y = 1.5 * a * j;
            p.push_back({x, y - 0.5 * a, 0});
            p.push_back({x, y + 0.5 * a, 0});
        }
    return p;
}

std::vector<P3> tetragoneSites(int nx, int ny, int nz, double a)
{
    const double s3 = std::sqrt(3.0);
    const P3 vi{a, 0, 0}, vj{a / 2, s3 / 2 * a, 0}, vk{a / 2, a / (2 * s3), std::sqrt(2.0) / s3 * a};
    std::vector<P3> p;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                p.push_back({i * vi[0] + j * vj[0] + k * vk[0], i * vi[1] + j * vj[1] + k * vk[1],
                             i * vi[2] + j * vj[2] + k * vk[2]});
    return p;
}

// the hexagonal comb sequence of the moire honeycomb flake: rings n = 0, 1, 2, ... walked through (m, l)
// (src/FHLattice.cpp:1356-1381)
void nextMoireComb(int& n, int& m, int& l)
{
    if (l == std::min(n, m + n)) {
        if (m == -n) {
            n = n + 1;
            m = n;
            l = 0;
        } else {
            m = m - 1;
            l = std::max(-n, m - n);
        }
    } else if (std::abs(m) == n) {
        l = l + 1;
    } else {
        l = std::min(n, m + n);
    }
}

P3 rotZ(const P3& v, double c, double s) { return {c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]}; }

std::vector<P3> moireHoneyCombSites(unsigned nCombs, double a, double b, double theta)
{
    // two flakes of nCombs combs, the second rotated by theta degrees about z and lifted by b.  The comb
    // lattice vectors are (sqrt3 a, 0) and (sqrt3 a / 2, 1.5): the second's y component is not scaled
    // by a in the reference (src/FHLattice.cpp:1191-1193), kept so that its files are reproduced.
    // Copies are removed from the second point of each layer on (:1229), with the L1 distance < 1e-3 a.
    const double s3 = std::sqrt(3.0), th = theta * kPi / 180;
    const std::array<P3, 6> c0 = combCorners(a);
    std::vector<P3> out;
    for (int layer = 0; layer < 2; ++layer) {
        const double c = std::cos(layer * th), s = std::sin(layer * th);
        const P3 g0 = rotZ({s3 * a, 0, 0}, c, s), g1 = rotZ({s3 * a / 2., 1.5, 0}, c, s);
        std::vector<P3> p;
        int n = 0, m = 0, l = 0;
        for (unsigned k = 0; k < nCombs; ++k) {
            const double u = m - l, v = l;
            const P3 shift{g0[0] * u + g1[0] * v, g0[1] * u + g1[1] * v, layer * b};
            for (const P3& q : c0) {
                const P3 r = rotZ(q, c, s);
                p.push_back({r[0] + shift[0], r[1] + shift[1], r[2] + shift[2]});
            }
            nextMoireComb(n, m, l);
        }
        removeCopies(p, 1, 1e-3 * a);
        out.insert(out.end(), p.begin(), p.end());
    }
    return out;
}

std::vector<P3> moireGrapheneSites(int nx, int ny, double a, double b, double theta)
{
    // two graphene sheets centred on the middle of their box, the second rotated by theta degrees and
    // placed at height b
    const double th = theta * kPi / 180;
    double box[9];
    grapheneBox(nx, ny, a, box);
    const double mx = (box[0] + box[1]) / 2, my = (box[3] + box[4]) / 2;
    std::vector<P3> out;
    for (int layer = 0; layer < 2; ++layer) {
        const double c = std::cos(layer * th), s = std::sin(layer * th);
        for (const P3& q : grapheneSites(nx, ny, a)) {
            P3 r = rotZ({q[0] - mx, q[1] - my, q[2]}, c, s);
            r[2] = layer * b;
            out.push_back(r);
        }
    }
    return out;
}

std::string fmtG(double x)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", x);  // the reference's stream default: 6 significant digits
    return buf;
}

}  // namespace

bool generateLattice(const std::string& typeName, const std::vector<double>& parameters, GeneratedLattice& out)
{
    std::vector<double> p = parameters;
    while (p.size() < 5) p.push_back(1.0);  // makeSitesSystem.cpp:128-131
    static const char* kNames[] = {"none", "grid", "bcc", "fcc", "triangular", "honeyCombs", "graphene", "moireHC",
                                   "moire", "kagome"};
    out = GeneratedLattice{};
    out.type = "none";
    for (const char* n : kNames)
        if (typeName == n) out.type = n;
    std::vector<P3> pos{{0, 0, 0}, {1, 0, 0}};  // types without a generator keep the two-site default
    std::fill(out.matrix, out.matrix + 9, 0.0);
    const std::string& t = out.type;
    if (t == "grid" || t == "bcc" || t == "fcc") {
        const unsigned nx = static_cast<unsigned>(static_cast<int>(p[0])), ny = static_cast<unsigned>(static_cast<int>(p[1])),
                       nz = static_cast<unsigned>(static_cast<int>(p[2]));
        const double a = p[3];
        pos = t == "grid" ? cubicSites(nx, ny, nz, a) : t == "bcc" ? bccSites(nx, ny, nz, a) : fccSites(nx, ny, nz, a);
        diagBox(nx, ny, nz, a, out.matrix);
    } else if (t == "honeyCombs") {
        const unsigned nc = static_cast<unsigned>(static_cast<int>(p[0]));
        pos = honeyCombSites(nc, p[1]);
        honeyCombBox(nc, p[1], out.matrix);
    } else if (t == "triangular") {
        const int nx = static_cast<int>(p[0]), ny = static_cast<int>(p[1]), nz = static_cast<int>(p[2]);
        pos = tetragoneSites(nx, ny, nz, p[3]);
        tetragoneBox(nx, ny, nz, p[3], out.matrix);
    } else if (t == "graphene") {
        const int nx = static_cast<int>(p[0]), ny = static_cast<int>(p[1]);
        pos = grapheneSites(nx, ny, p[2]);
        grapheneBox(nx, ny, p[2], out.matrix);
    } else if (t == "moireHC") {
        const unsigned nc = static_cast<unsigned>(static_cast<int>(p[0]));
        pos = moireHoneyCombSites(nc, p[1], p[2], p[3]);
        honeyCombBox(nc, p[1], out.matrix);
    } else if (t == "moire") {
        const int nx = static_cast<int>(p[0]), ny = static_cast<int>(p[1]);
        pos = moireGrapheneSites(nx, ny, p[2], p[3], p[4]);
        grapheneBox(nx, ny, p[2], out.matrix);
    }
    out.positions.clear();
    for (const P3& q : pos) out.positions.insert(out.positions.end(), q.begin(), q.end());
    return true;
}

bool writeXyzFile(const std::string& path, const std::vector<double>& positions, const std::string& siteName)
{
    // FHLattice::writePositions (src/FHLattice.cpp:715-738): the count, an empty comment line, then one
    // tab-separated "name x y z" line per site, without a final newline
    std::ofstream f(path);
    if (!f.is_open()) return false;
    const size_t n = positions.size() / 3;
    f << n << "\n";
    for (size_t i = 0; i < n; ++i)
        f << "\n" << siteName << "\t" << fmtG(positions[3 * i]) << "\t" << fmtG(positions[3 * i + 1]) << "\t"
          << fmtG(positions[3 * i + 2]);
    return f.good();
}

fpsjson::Value sitesSystemJson(const std::string& name, const GeneratedLattice& lat, const std::string& ic)
{
    using fpsjson::Value;
    Value sim = Value::object();
    sim["sim_name"] = name;
    sim["method"] = "HF";
    Value par = Value::object();
    par["t_max"] = 2.0;
    par["dt"] = 0.01;
    par["stepsPerTimePoints"] = 1;
    par["n_traj"] = 1000;
    par["int_order"] = 4;
    sim["parameters"] = par;

    Value sys = Value::object();
    Value l = Value::object();
    l["latticeType"] = lat.type;
    l["positionsFile"] = name + ".xyz";
    l["periodic"] = false;
    Value mat = Value::array();
    for (int i = 0; i < 3; ++i) {
        Value row = Value::array();
        for (int j = 0; j < 3; ++j) row.append(lat.matrix[3 * i + j]);
        mat.append(row);
    }
    l["latticeMatrix"] = mat;
    Value zero = Value::array();
    for (int j = 0; j < 3; ++j) zero.append(0.0);
    l["latticeZero"] = zero;
    sys["lattice"] = l;
    Value icj = Value::object();
    if (ic == "chargeWave") {
        icj["type"] = "givenDiagonal";
        icj["subtype"] = "chargeWave";
    } else if (ic == "tbGroundState" || ic == "mfGroundState") {
        icj["type"] = ic;
    } else {
        icj["type"] = "givenDiagonal";
        icj["subtype"] = "spinWave";
    }
    sys["IC"] = icj;
    // setBasisInteractions: nearest neighbours at distance 1 with t = -1, on-site U = 1
    Value tun = Value::object();
    tun["type"] = "nearestNeighbors";
    tun["neighborDistance"] = 1.0;
    tun["neighborInteraction"] = -1.0;
    sys["tunnelingInteraction"] = tun;
    Value cou = Value::object();
    cou["type"] = "onSite";
    cou["onSiteInteraction"] = 1.0;
    sys["coulombInteraction"] = cou;
    Value ext = Value::object();
    ext["type"] = "null";
    sys["externalPotential"] = ext;
    sim["system"] = sys;

    Value res = Value::object();
    res["diagDMatrix"] = true;
    res["diag2DMatrix"] = false;
    res["onSiteDoublons"] = false;
    res["wholeDMatrix"] = false;
    res["tbStates"] = false;
    res["mfStates"] = false;
    res["energies"] = true;
    sim["results"] = res;
    return sim;
}

}  // namespace fpsamd
