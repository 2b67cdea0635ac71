// eigen_ref.cpp -- the eigen step of WPCA computed with the reference's own
// solver: Eigen::SelfAdjointEigenSolver<Matrix3f>::computeDirect from the Eigen
// headers vendored in the reference (vendor/eigen/include), followed by the
// clamp and G = ParticleRadiusInv * R * diag(1/Sigma) * R^T as
// RayMarcher.cpp:179-253 evaluates them, and glm's mat3 determinant.
//
// TEST INFRASTRUCTURE ONLY: oracle/Makefile builds it into oracle/_ref/ when
// the reference tree is present; only tests load it, to pin the oracle's
// Jacobi eigen-decomposition against the reference's solver.
#include <Eigen/Dense>

#include <algorithm>
#include <cstdint>

extern "C" {

// Eigenvalues (ascending, as computeDirect returns them) and eigenvectors
// (column k of V, row major) of the symmetric matrix with lower triangle
// c = (c00, c10, c11, c20, c21, c22).
int ref_eig3_direct(const float c[6], float lam[3], float V[9])
{
    Eigen::Matrix3f C;
    C << c[0], c[1], c[3],
         c[1], c[2], c[4],
         c[3], c[4], c[5];
    Eigen::SelfAdjointEigenSolver<Eigen::Matrix3f> solver;
    solver.computeDirect(C);
    if (solver.info() != Eigen::Success) return -1;
    const Eigen::Vector3f S = solver.eigenvalues();
    const Eigen::Matrix3f R = solver.eigenvectors();
    for (int k = 0; k < 3; k++) lam[k] = S[k];
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) V[3 * r + k] = R(r, k);
    return 0;
}

// G (row major) and det G from C and the neighbour count n, with the
// reference's solver, clamp and product.
int ref_aniso_G_from_C(const float c[6], uint32_t n, float hinv, float k_n, float k_r, float k_s, int32_t n_eps,
                       float G[9], float *detG)
{
    Eigen::Matrix3f C;
    C << c[0], c[1], c[3],
         c[1], c[2], c[4],
         c[3], c[4], c[5];
    Eigen::SelfAdjointEigenSolver<Eigen::Matrix3f> solver;
    solver.computeDirect(C);
    if (solver.info() != Eigen::Success) return -1;
    Eigen::Vector3f S = solver.eigenvalues();
    const Eigen::Matrix3f R = solver.eigenvectors();
    if (n <= (uint32_t)n_eps) {
        S.setConstant(k_n);
    } else {
        const float mn = std::max(S[0], std::max(S[1], S[2])) / k_r;
        for (int k = 0; k < 3; k++) S[k] = std::max(S[k], mn);
        S *= k_s;
    }
    const Eigen::Matrix3f Gm = hinv * R * S.cwiseInverse().asDiagonal() * R.transpose();
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) G[3 * r + k] = Gm(r, k);
    // glm::determinant of the glm matrix whose column k is Gm.col(k): m[k][r] = Gm(r, k)
    auto m = [&](int col, int row) { return Gm(row, col); };
    *detG = +m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2))
            - m(1, 0) * (m(0, 1) * m(2, 2) - m(2, 1) * m(0, 2))
            + m(2, 0) * (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2));
    return 0;
}

}  // extern "C"
