// eigen.cpp -- host eigen-decomposition for the eigenbasis initial states.
//
// The reference builds tbGroundState / mfGroundState initial conditions from LAPACK's eig_sym
// (src/FHSystem.cpp:911-1019): the rotation R is the eigenvector matrix of the (mean-field) hopping
// and the lowest nPart eigenstates are occupied.  No LAPACK is linked here, so the decomposition is a
// cyclic Jacobi method on the real symmetric matrix (these matrices are real in every system the JSON
// front end builds): rotations zero the off-diagonal entries sweep by sweep until their norm is below
// 1e-14 of the matrix norm, then one more sweep.  Eigenvalues come out ascending; each eigenvector carries a fixed sign
// (its largest-magnitude entry, the first one within 1e-8 on ties, is positive), so the basis is reproducible.
// Within a degenerate eigenspace any orthonormal basis is as valid as LAPACK's; only occupations that
// split a degenerate level depend on that choice (as they do in the reference).
#include "eigen.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace fpsamd {

int symmetric_eigen(int n, const double* a, double* evals, double* evecs)
{
    if (n <= 0) return 0;
    std::vector<double> A(a, a + static_cast<size_t>(n) * n);
    std::vector<double> V(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[static_cast<size_t>(i) * n + i] = 1.0;
    auto at = [&](std::vector<double>& M, int i, int j) -> double& { return M[static_cast<size_t>(i) * n + j]; };
    double norm = 0.0;
    for (double x : A) norm += x * x;
    norm = std::sqrt(norm);
    const double tol = 1e-14 * (norm > 0 ? norm : 1.0);
    int sweep = 0, polish = 0;
    for (; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += at(A, i, j) * at(A, i, j);
        // one more sweep after the threshold: convergence is quadratic, so the eigenvectors end
        // accurate to rounding even across close eigenvalues
        if (std::sqrt(2.0 * off) <= tol && polish++ > 0) break;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double apq = at(A, p, q);
                if (std::fabs(apq) <= 1e-300) continue;
                const double app = at(A, p, p), aqq = at(A, q, q);
                // rotation angle that zeroes A_pq (Golub & Van Loan, Algorithm 8.4.1)
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
                    const double akp = at(A, k, p), akq = at(A, k, q);
                    at(A, k, p) = c * akp - s * akq;
                    at(A, k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = at(A, p, k), aqk = at(A, q, k);
                    at(A, p, k) = c * apk - s * aqk;
                    at(A, q, k) = s * apk + c * aqk;
                }
                at(A, p, q) = 0.0;
                at(A, q, p) = 0.0;
                for (int k = 0; k < n; ++k) {  // V <- V J
                    const double vkp = at(V, k, p), vkq = at(V, k, q);
                    at(V, k, p) = c * vkp - s * vkq;
                    at(V, k, q) = s * vkp + c * vkq;
                }
            }
        }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return at(A, x, x) < at(A, y, y); });
    for (int c = 0; c < n; ++c) {
        const int src = order[c];
        evals[c] = at(A, src, src);
        int big = 0;
        for (int k = 1; k < n; ++k)
            if (std::fabs(at(V, k, src)) > std::fabs(at(V, big, src)) * (1.0 + 1e-8)) big = k;
        const double sgn = at(V, big, src) < 0 ? -1.0 : 1.0;
        for (int k = 0; k < n; ++k) evecs[static_cast<size_t>(k) * n + c] = sgn * at(V, k, src);
    }
    return sweep < 100 ? 0 : -1;
}

int ground_state_rotation(int L, const double* tM, const double* U, int n_up, int n_dn, bool mean_field,
                          double* rot_up, double* rot_dn, int* iterations, double* residue, double* e_up,
                          double* e_dn)
{
    std::vector<double> ev(L), ev_dn(L), M(static_cast<size_t>(L) * L);
    if (iterations) *iterations = 0;
    if (residue) *residue = 0.0;
    if (!mean_field) {  // src/FHSystem.cpp:911-937: eig_sym(tM), both spins share the basis
        if (symmetric_eigen(L, tM, ev.data(), rot_up) != 0) return -1;
        std::copy(rot_up, rot_up + static_cast<size_t>(L) * L, rot_dn);
        if (e_up) std::copy(ev.begin(), ev.end(), e_up);
        if (e_dn) std::copy(ev.begin(), ev.end(), e_dn);
        return 0;
    }
    // src/FHSystem.cpp:940-1019: Hartree fields from the other spin's density, eig_sym per spin,
    // densities from the lowest nPart states, plain substitution until the residue < 1e-4 L, at most
    // 100 iterations.  The starting densities are zero (an Armadillo cx_dvec(n) is zero-filled).
    std::vector<double> nu(L, 0.0), nd(L, 0.0), nu_new(L), nd_new(L);
    const double threshold = 1e-4 * L;
    int it = 0;
    double res = 0.0;
    auto density = [&](const double* V, int npart, std::vector<double>& out) {
        for (int i = 0; i < L; ++i) {
            double s = 0.0;
            for (int k = 0; k < npart && k < L; ++k) s += V[static_cast<size_t>(i) * L + k] * V[static_cast<size_t>(i) * L + k];
            out[i] = s;
        }
    };
    while (it < 100) {
        for (int spin = 0; spin < 2; ++spin) {
            const std::vector<double>& other = spin == 0 ? nd : nu;
            std::copy(tM, tM + static_cast<size_t>(L) * L, M.begin());
            for (int i = 0; i < L; ++i) {
                double h = 0.0;
                for (int j = 0; j < L; ++j) h += (U ? U[static_cast<size_t>(i) * L + j] : 0.0) * other[j];
                M[static_cast<size_t>(i) * L + i] += h;
            }
            if (symmetric_eigen(L, M.data(), spin == 0 ? ev.data() : ev_dn.data(), spin == 0 ? rot_up : rot_dn) != 0)
                return -1;
        }
        density(rot_up, n_up, nu_new);
        density(rot_dn, n_dn, nd_new);
        ++it;
        double r1 = 0.0, r2 = 0.0;
        for (int i = 0; i < L; ++i) {
            r1 += (nu_new[i] - nu[i]) * (nu_new[i] - nu[i]);
            r2 += (nd_new[i] - nd[i]) * (nd_new[i] - nd[i]);
        }
        res = std::sqrt(r1) + std::sqrt(r2);
        if (res < threshold) break;
        nu = nu_new;
        nd = nd_new;
    }
    if (iterations) *iterations = it;
    if (residue) *residue = res;
    if (e_up) std::copy(ev.begin(), ev.end(), e_up);
    if (e_dn) std::copy(ev_dn.begin(), ev_dn.end(), e_dn);
    return 0;
}

}  // namespace fpsamd
