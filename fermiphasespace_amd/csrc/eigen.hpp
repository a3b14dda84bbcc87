// eigen.hpp -- host eigen-decomposition for the eigenbasis initial states (eigen.cpp).
#pragma once

namespace fpsamd {

// Eigenvalues (ascending) and eigenvectors (columns, row-major evecs[k*n + c]) of the real symmetric
// n x n matrix a.  Each eigenvector's largest-magnitude entry (the first within 1e-8
// on ties) is positive.  0 on success.
int symmetric_eigen(int n, const double* a, double* evals, double* evecs);

// Rotation matrices of the tbGroundState (mean_field false) or mfGroundState (true) initial states:
// eigenvector columns of tM, or of tM + diag(U n_other) iterated to self-consistency.  rot_* are
// row-major [L][L]; the lowest n_up / n_dn columns are the occupied states.  e_up / e_dn (may be null)
// receive the eigenvalues of each spin's basis in ascending order (both tM's for tight binding; the last
// iteration's for mean field).  0 on success.
int ground_state_rotation(int L, const double* tM, const double* U, int n_up, int n_dn, bool mean_field,
                          double* rot_up, double* rot_dn, int* iterations, double* residue, double* e_up = nullptr,
                          double* e_dn = nullptr);

}  // namespace fpsamd
