/*
 * zmpm_oracle.cpp -- FP64 CPU restatement of Ziran's explicit APIC MPM substep
 * and AnisoMPM explicit damage update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against.  Only tests/, the smoke check and the bench's CPU-baseline
 * leg may load it; the product library never links or calls it.
 *
 * It implements the same C ABI as the GPU library (include/zmpm.h) so a test
 * can drive both with identical calls, plus a few zmpm_oracle_* unit entry
 * points used to pin individual pieces against closed forms.
 *
 * Faithfulness: every routine cites the reference lines it restates.  The
 * reference cannot be built here (Eigen3 / TBB headers and OpenVDB are absent),
 * so bit-level agreement with Eigen's evaluation order is not claimed; the
 * algorithms, operation order inside each routine, accumulation order of P2G
 * (8-colour block schedule, sorted particle order) and all branch conventions
 * follow the reference.  Arithmetic is double.  When the handle is created in
 * the FP32 layout (4x4x4 blocks), positions are held on the GPU library's FP32
 * position grid -- x / dx in units of 2^-21, rounded at upload and after every
 * advection -- and the base node is the exact floor(x / dx - 1/2) of that
 * fixed-point value, as the FP32 form computes it, so keys can be compared
 * bit-exactly.
 *
 * Compiled with -ffp-contract=off: the reference's base-node flooring must not
 * be FMA-contracted (MpmSimulationBase.cpp:1545).
 */
#include "zmpm.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <omp.h>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------- 3x3 algebra
struct M3 {
    double a[3][3];
};

inline M3 mzero()
{
    M3 r;
    std::memset(&r, 0, sizeof(r));
    return r;
}
inline M3 meye()
{
    M3 r = mzero();
    r.a[0][0] = r.a[1][1] = r.a[2][2] = 1;
    return r;
}
inline M3 mmul(const M3& A, const M3& B)
{
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
    return r;
}
inline M3 mtrans(const M3& A)
{
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.a[i][j] = A.a[j][i];
    return r;
}
inline M3 mload(const double* p)
{
    M3 r;
    for (int i = 0; i < 9; i++)
        r.a[i / 3][i % 3] = p[i];
    return r;
}
inline void mstore(const M3& A, double* p)
{
    for (int i = 0; i < 9; i++)
        p[i] = A.a[i / 3][i % 3];
}

// Eigen's 3x3 determinant expansion along the first column
// (used by F.determinant() at CorotatedElasticity.cpp:60, NeoHookean.cpp:27).
inline double mdet(const M3& m)
{
    const double(*a)[3] = m.a;
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
        - a[1][0] * (a[0][1] * a[2][2] - a[0][2] * a[2][1])
        + a[2][0] * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
}

// cofactorMatrix: J F^{-T}  (Lib/Ziran/Math/Linear/DenseExt.h:240-252)
inline M3 mcofactor(const M3& Fm)
{
    const double(*F)[3] = Fm.a;
    M3 A;
    A.a[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    A.a[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    A.a[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    A.a[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    A.a[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    A.a[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    A.a[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    A.a[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    A.a[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    return A;
}

// 3x3 inverse through the cofactor matrix (Eigen's compute_inverse for size 3).
inline M3 minverse(const M3& A)
{
    M3 cof = mcofactor(A); // cof = det * A^{-T}
    double det = cof.a[0][0] * A.a[0][0] + cof.a[0][1] * A.a[0][1] + cof.a[0][2] * A.a[0][2];
    double inv = 1.0 / det;
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.a[i][j] = cof.a[j][i] * inv;
    return r;
}

// ------------------------------------------------------- Givens / implicit QR SVD
// Givens rotation with the conventions of Lib/Ziran/Math/Linear/Givens.h:32-181.
struct Givens {
    int i, k;
    double c, s;
    Givens(int i_, int k_)
        : i(i_), k(k_), c(1), s(0)
    {
    }
    // Givens.h:72-85
    void compute(double a, double b)
    {
        double d = a * a + b * b;
        c = 1;
        s = 0;
        double sqrtd = std::sqrt(d);
        if (sqrtd) {
            double t = 1 / sqrtd;
            c = a * t;
            s = -b * t;
        }
    }
    // Givens.h:92-105
    void computeUnconventional(double a, double b)
    {
        double d = a * a + b * b;
        c = 0;
        s = 1;
        double sqrtd = std::sqrt(d);
        if (sqrtd) {
            double t = 1 / sqrtd;
            s = a * t;
            c = b * t;
        }
    }
    // Givens.h:131-140 (rows i,k of A)
    void rowRotation(M3& A) const
    {
        for (int j = 0; j < 3; j++) {
            double t1 = A.a[i][j], t2 = A.a[k][j];
            A.a[i][j] = c * t1 - s * t2;
            A.a[k][j] = s * t1 + c * t2;
        }
    }
    // Givens.h:150-159 (columns i,k of A)
    void columnRotation(M3& A) const
    {
        for (int j = 0; j < 3; j++) {
            double t1 = A.a[j][i], t2 = A.a[j][k];
            A.a[j][i] = c * t1 - s * t2;
            A.a[j][k] = s * t1 + c * t2;
        }
    }
    // Givens.h:164-170
    void mulAssign(const Givens& B)
    {
        double nc = c * B.c - s * B.s;
        double ns = s * B.c + c * B.s;
        c = nc;
        s = ns;
    }
};

// zeroChase (Givens.h:195-240)
void zeroChase(M3& H, M3& U, M3& V)
{
    Givens r1(0, 1);
    r1.compute(H.a[0][0], H.a[1][0]);
    Givens r2(1, 2);
    if (H.a[1][0] != 0)
        r2.compute(H.a[0][0] * H.a[0][1] + H.a[1][0] * H.a[1][1], H.a[0][0] * H.a[0][2] + H.a[1][0] * H.a[1][2]);
    else
        r2.compute(H.a[0][1], H.a[0][2]);
    r1.rowRotation(H);
    r2.columnRotation(H);
    r2.columnRotation(V);
    Givens r3(1, 2);
    r3.compute(H.a[1][1], H.a[2][1]);
    r3.rowRotation(H);
    r1.columnRotation(U);
    r3.columnRotation(U);
}

// makeUpperBidiag (Givens.h:252-271)
void makeUpperBidiag(M3& H, M3& U, M3& V)
{
    U = meye();
    V = meye();
    Givens r(1, 2);
    r.compute(H.a[1][0], H.a[2][0]);
    r.rowRotation(H);
    r.columnRotation(U);
    zeroChase(H, U, V);
}

// wilkinsonShift (ImplicitQRSVD.h:260-272)
double wilkinsonShift(double a1, double b1, double a2)
{
    double d = 0.5 * (a1 - a2);
    double bs = b1 * b1;
    return a2 - std::copysign(bs / (std::fabs(d) + std::sqrt(d * d + bs)), d);
}

// 2x2 SVD of the block B[t..t+1][t..t+1] in Givens form
// (polarDecomposition 2x2 ImplicitQRSVD.h:45-66, singularValueDecomposition 2x2 :116-175)
void svd2x2(const M3& B, int t, Givens& U, double& s0, double& s1, Givens& V)
{
    double A00 = B.a[t][t], A01 = B.a[t][t + 1], A10 = B.a[t + 1][t], A11 = B.a[t + 1][t + 1];
    double x0 = A00 + A11, x1 = A10 - A01;
    double denominator = std::sqrt(x0 * x0 + x1 * x1);
    U.c = 1;
    U.s = 0;
    if (denominator != 0) {
        U.c = x0 / denominator;
        U.s = -x1 / denominator;
    }
    // S = A; U.rowRotation(S) on a 2x2
    double S00 = U.c * A00 - U.s * A10;
    double S01 = U.c * A01 - U.s * A11;
    double S11 = U.s * A01 + U.c * A11;
    double cosine, sine;
    double x = S00, y = S01, z = S11;
    double y2 = y * y;
    if (y2 == 0) {
        cosine = 1;
        sine = 0;
        s0 = x;
        s1 = z;
    }
    else {
        double tau = 0.5 * (x - z);
        double w = std::sqrt(tau * tau + y2);
        double tt;
        if (tau > 0)
            tt = y / (tau + w);
        else
            tt = y / (tau - w);
        cosine = 1 / std::sqrt(tt * tt + 1);
        sine = -tt * cosine;
        double c2 = cosine * cosine;
        double csy = 2 * cosine * sine * y;
        double s2 = sine * sine;
        s0 = c2 * x - csy + s2 * z;
        s1 = s2 * x + csy + c2 * z;
    }
    if (s0 < s1) {
        std::swap(s0, s1);
        V.c = -sine;
        V.s = cosine;
    }
    else {
        V.c = cosine;
        V.s = sine;
    }
    U.mulAssign(V);
}

// process<t> (ImplicitQRSVD.h:277-293)
void process(int t, M3& B, M3& U, double* sigma, M3& V)
{
    int other = (t == 1) ? 0 : 2;
    Givens u(0, 1), v(0, 1);
    sigma[other] = B.a[other][other];
    svd2x2(B, t, u, sigma[t], sigma[t + 1], v);
    u.i += t;
    u.k += t;
    v.i += t;
    v.k += t;
    u.columnRotation(U);
    v.columnRotation(V);
}

// flipSign (ImplicitQRSVD.h:298-303)
void flipSign(int i, M3& U, double* sigma)
{
    sigma[i] = -sigma[i];
    for (int r = 0; r < 3; r++)
        U.a[r][i] = -U.a[r][i];
}
void swapCols(M3& A, int i, int j)
{
    for (int r = 0; r < 3; r++)
        std::swap(A.a[r][i], A.a[r][j]);
}
void negCol(M3& A, int i)
{
    for (int r = 0; r < 3; r++)
        A.a[r][i] = -A.a[r][i];
}

// sort<0> (ImplicitQRSVD.h:308-341)
void sort0(M3& U, double* sigma, M3& V)
{
    if (std::fabs(sigma[1]) >= std::fabs(sigma[2])) {
        if (sigma[1] < 0) {
            flipSign(1, U, sigma);
            flipSign(2, U, sigma);
        }
        return;
    }
    if (sigma[2] < 0) {
        flipSign(1, U, sigma);
        flipSign(2, U, sigma);
    }
    std::swap(sigma[1], sigma[2]);
    swapCols(U, 1, 2);
    swapCols(V, 1, 2);
    if (sigma[1] > sigma[0]) {
        std::swap(sigma[0], sigma[1]);
        swapCols(U, 0, 1);
        swapCols(V, 0, 1);
    }
    else {
        negCol(U, 2);
        negCol(V, 2);
    }
}

// sort<1> (ImplicitQRSVD.h:307-345 second overload)
void sort1(M3& U, double* sigma, M3& V)
{
    if (std::fabs(sigma[0]) >= sigma[1]) {
        if (sigma[0] < 0) {
            flipSign(0, U, sigma);
            flipSign(2, U, sigma);
        }
        return;
    }
    std::swap(sigma[0], sigma[1]);
    swapCols(U, 0, 1);
    swapCols(V, 0, 1);
    if (std::fabs(sigma[1]) < std::fabs(sigma[2])) {
        std::swap(sigma[1], sigma[2]);
        swapCols(U, 1, 2);
        swapCols(V, 1, 2);
    }
    else {
        negCol(U, 1);
        negCol(V, 1);
    }
    if (sigma[1] < 0) {
        flipSign(1, U, sigma);
        flipSign(2, U, sigma);
    }
}

// singularValueDecompositionHelper (ImplicitQRSVD.h:354-517): the implicit-shift
// symmetric QR SVD of Gast, Fu, Jiang and Teran (UCLA CAM 16-04, 2016).
int svd3(const M3& A, M3& U, double* sigma, M3& V, double tol)
{
    M3 B = A;
    U = meye();
    V = meye();
    makeUpperBidiag(B, U, V);
    int count = 0;
    double mu = 0;
    Givens r(0, 1);
    double alpha_1 = B.a[0][0], beta_1 = B.a[0][1], alpha_2 = B.a[1][1], alpha_3 = B.a[2][2], beta_2 = B.a[1][2];
    double gamma_1 = alpha_1 * beta_1, gamma_2 = alpha_2 * beta_2;
    tol *= std::max(0.5 * std::sqrt(alpha_1 * alpha_1 + alpha_2 * alpha_2 + alpha_3 * alpha_3 + beta_1 * beta_1 + beta_2 * beta_2), 1.0);
    while (std::fabs(beta_2) > tol && std::fabs(beta_1) > tol && std::fabs(alpha_1) > tol
        && std::fabs(alpha_2) > tol && std::fabs(alpha_3) > tol) {
        mu = wilkinsonShift(alpha_2 * alpha_2 + beta_1 * beta_1, gamma_2, alpha_3 * alpha_3 + beta_2 * beta_2);
        r.compute(alpha_1 * alpha_1 - mu, gamma_1);
        r.columnRotation(B);
        r.columnRotation(V);
        zeroChase(B, U, V);
        alpha_1 = B.a[0][0];
        beta_1 = B.a[0][1];
        alpha_2 = B.a[1][1];
        alpha_3 = B.a[2][2];
        beta_2 = B.a[1][2];
        gamma_1 = alpha_1 * beta_1;
        gamma_2 = alpha_2 * beta_2;
        count++;
    }
    if (std::fabs(beta_2) <= tol) {
        process(0, B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else if (std::fabs(beta_1) <= tol) {
        process(1, B, U, sigma, V);
        sort1(U, sigma, V);
    }
    else if (std::fabs(alpha_2) <= tol) {
        Givens r1(1, 2);
        r1.computeUnconventional(B.a[1][2], B.a[2][2]);
        r1.rowRotation(B);
        r1.columnRotation(U);
        process(0, B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else if (std::fabs(alpha_3) <= tol) {
        Givens r1(1, 2);
        r1.compute(B.a[1][1], B.a[1][2]);
        r1.columnRotation(B);
        r1.columnRotation(V);
        Givens r2(0, 2);
        r2.compute(B.a[0][0], B.a[0][2]);
        r2.columnRotation(B);
        r2.columnRotation(V);
        process(0, B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else if (std::fabs(alpha_1) <= tol) {
        Givens r1(0, 1);
        r1.computeUnconventional(B.a[0][1], B.a[1][1]);
        r1.rowRotation(B);
        r1.columnRotation(U);
        Givens r2(0, 2);
        r2.computeUnconventional(B.a[0][2], B.a[2][2]);
        r2.rowRotation(B);
        r2.columnRotation(U);
        process(1, B, U, sigma, V);
        sort1(U, sigma, V);
    }
    return count;
}

const double kSvdTol = 1024 * std::numeric_limits<double>::epsilon(); // ImplicitQRSVD.h:526-533
constexpr int kDamageCgIterations = 100; // ConjugateGradient<..>(100) (AnisoFractureSimulation.h:70)

// polarDecomposition rotation R = U V^T (ImplicitQRSVD.h:546-558)
M3 polarR(const M3& F)
{
    M3 U, V;
    double s[3];
    svd3(F, U, s, V, kSvdTol);
    return mmul(U, mtrans(V));
}

// GivensQR (GivensQR.h:20-48): R <- G_n..G_1 A, Q = (G_n..G_1)^T
void givensQR(const M3& A, M3& Q, M3& R)
{
    R = A;
    Q = meye();
    for (int j = 0; j < 3; j++)
        for (int i = 2; i > j; i--) {
            Givens r(i - 1, i);
            r.compute(R.a[i - 1][j], R.a[i][j]);
            r.rowRotation(R);
            r.rowRotation(Q);
        }
    Q = mtrans(Q);
}

// -------------------------------------------------------------- constitutive
struct Material {
    zmpm_material p;
};

// QRStableNeoHookean::updateScratch + firstPiola (QRStableNeoHookean.h:108-213,
// QrBasedHelper.h:28-40).  Returns P; J and energy through pointers.
M3 qrStablePiola(const zmpm_material& mat, double g, double kx, double ky, const M3& F, double* J_out, double* psi_out)
{
    const double mu = mat.mu, lambda = mat.lambda;
    M3 Q, R;
    givensQR(F, Q, R);
    if (R.a[0][0] < 0) { // sign convention, QRStableNeoHookean.h:125-139
        R.a[0][0] *= -1;
        for (int r = 0; r < 3; r++) Q.a[r][0] *= -1;
        R.a[0][1] *= -1;
        R.a[0][2] *= -1;
        R.a[2][2] *= -1;
        for (int r = 0; r < 3; r++) Q.a[r][2] *= -1;
    }
    if (R.a[1][1] < 0) {
        R.a[1][1] *= -1;
        for (int r = 0; r < 3; r++) Q.a[r][1] *= -1;
        R.a[1][2] *= -1;
        R.a[2][2] *= -1;
        for (int r = 0; r < 3; r++) Q.a[r][2] *= -1;
    }
    M3 Rinv = minverse(R);
    double J = 1;
    for (int i = 0; i < 3; i++) J *= R.a[i][i];
    auto sqr = [](double v) { return v * v; };
    double energy = 0;
    double vol = (J >= 1) ? g * 0.5 * lambda * sqr(J - 1) : 0.5 * lambda * sqr(J - 1);
    energy += vol;
    double rss = 0;
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) rss += sqr(R.a[i][j]);
    energy += g * (mu / 2 * (rss - 3) - mu * (J - 1));
    energy += kx / 2 * sqr(R.a[0][0] - 1);
    energy += ky / 2 * sqr(std::sqrt(sqr(R.a[0][1]) + sqr(R.a[1][1])) - 1);
    M3 D = mzero(); // dPsiHat/dR
    if (J >= 1)
        for (int i = 0; i < 3; i++) D.a[i][i] += g * lambda * (J - 1) * J / R.a[i][i];
    else
        for (int i = 0; i < 3; i++) D.a[i][i] += lambda * (J - 1) * J / R.a[i][i];
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) D.a[i][j] += g * mu * R.a[i][j];
    for (int i = 0; i < 3; i++) D.a[i][i] -= g * mu * J / R.a[i][i];
    D.a[0][0] += kx * (R.a[0][0] - 1);
    {
        double ss = std::sqrt(sqr(R.a[0][1]) + sqr(R.a[1][1]));
        D.a[0][1] += ky * (ss - 1) * R.a[0][1] / ss;
        D.a[1][1] += ky * (ss - 1) * R.a[1][1] / ss;
    }
    // evaluateP: A = D R^T, symmetrised from its upper triangle; P = Q A R^{-T}
    M3 A = mmul(D, mtrans(R));
    for (int i = 1; i < 3; i++)
        for (int j = 0; j < i; j++) A.a[i][j] = A.a[j][i];
    if (J_out) *J_out = J;
    if (psi_out) *psi_out = energy;
    return mmul(mmul(Q, A), mtrans(Rinv));
}

// QRAnisotropic::updateScratch + firstPiola (QRAnisotropic.h:87-169,
// QrBasedHelper.h:28-40) with k_shear = mu, k_volume = lambda and fibre
// stiffnesses kf (QRAnisotropic.h:57-64).  No sign convention on R: the Givens
// sweep leaves R_00, R_11 >= 0 and R_22 with the sign of det F.  Returns P; J
// and energy through pointers.
M3 qrAnisoPiola(const zmpm_material& mat, double g, const double kf[3], const M3& F, double* J_out, double* psi_out)
{
    const double kv = mat.lambda, ks = mat.mu;
    const bool iso = mat.isotropic != 0;
    M3 Q, R;
    givensQR(F, Q, R);
    M3 Rinv = minverse(R);
    double J = 1;
    for (int i = 0; i < 3; i++) J *= R.a[i][i];
    auto sqr = [](double v) { return v * v; };
    double energy = (J >= 1) ? g * 0.5 * kv * sqr(J - 1) : 0.5 * kv * sqr(J - 1);
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) energy += g * 0.5 * ks * sqr(R.a[i][j]);
    for (int i = 0; i < 3; i++) {
        const double gi = (i == 0 && !iso) ? 1.0 : g; // the first fibre is not degraded
        energy += gi * 0.5 * kf[i] * sqr(R.a[i][i] - 1);
    }
    M3 D = mzero(); // dPsiHat/dR
    for (int i = 0; i < 3; i++) D.a[i][i] += (J >= 1 ? g : 1.0) * kv * (J - 1) * J / R.a[i][i];
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) D.a[i][j] += g * ks * R.a[i][j];
    for (int i = 0; i < 3; i++) {
        const double gi = (i == 0 && !iso) ? 1.0 : g;
        D.a[i][i] += gi * kf[i] * (R.a[i][i] - 1);
    }
    M3 A = mmul(D, mtrans(R));
    for (int i = 1; i < 3; i++)
        for (int j = 0; j < i; j++) A.a[i][j] = A.a[j][i];
    if (J_out) *J_out = J;
    if (psi_out) *psi_out = energy;
    return mmul(mmul(Q, A), mtrans(Rinv));
}

// NeoHookeanBorden::kirchhoff (NeoHookeanBorden.cpp:78-92) with kappa =
// 0.666666 mu + lambda (setLameParameters :15-21): tau = mu J^{-2/3} dev(F F^T)
// + J kappa/2 (J - 1/J) I, the volumetric part degraded only for J >= 1.
double bordenKappa(const zmpm_material& mat) { return 0.666666 * mat.mu + mat.lambda; }

M3 bordenTau(const zmpm_material& mat, const M3& F, double g)
{
    const double J = mdet(F);
    M3 B = mmul(F, mtrans(F));
    const double tr3 = (B.a[0][0] + B.a[1][1] + B.a[2][2]) / 3.0;
    const double sdev = mat.mu * std::pow(J, -2.0 / 3.0);
    const double vol = J * (bordenKappa(mat) / 2 * (J - 1 / J));
    const double gd = g, gv = (J >= 1) ? g : 1.0;
    M3 tau;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            tau.a[i][j] = gd * sdev * (B.a[i][j] - (i == j ? tr3 : 0.0)) + (i == j ? gv * vol : 0.0);
    return tau;
}

// NeoHookeanBorden::psi (NeoHookeanBorden.cpp:45-59)
double bordenPsi(const zmpm_material& mat, const M3& F, double g)
{
    const double J = mdet(F);
    const double a = std::pow(J, -1.0 / 3.0);
    double trC = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) trC += (a * F.a[i][j]) * (a * F.a[i][j]);
    const double dev = mat.mu / 2 * (trC - 3);
    const double vol = bordenKappa(mat) / 2 * ((J * J - 1) / 2 - std::log(J));
    return (J >= 1) ? g * (vol + dev) : g * dev + vol;
}

// Models that carry the phase field (updateDamage, AnisoFractureSimulation.h:342-363).
bool phaseFieldModel(int model)
{
    return model == ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN || model == ZMPM_MODEL_QR_ANISOTROPIC
        || model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN;
}

// First Piola stress of a material at F with degradation g.
M3 firstPiola(const zmpm_material& mat, const M3& F, double g)
{
    switch (mat.model) {
    case ZMPM_MODEL_FIXED_COROTATED: {
        // CorotatedElasticity::updateScratch / firstPiola (CorotatedElasticity.cpp:55-76)
        double J = mdet(F);
        M3 JFinvT = mcofactor(F);
        M3 R = polarR(F);
        M3 P;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                P.a[i][j] = 2 * mat.mu * (F.a[i][j] - R.a[i][j]) + mat.lambda * (J - 1) * JFinvT.a[i][j];
        return P;
    }
    case ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN:
        return qrStablePiola(mat, g, mat.kx, mat.ky, F, nullptr, nullptr);
    case ZMPM_MODEL_QR_ANISOTROPIC:
        return qrAnisoPiola(mat, g, mat.fiber_stiffness, F, nullptr, nullptr);
    case ZMPM_MODEL_NEO_HOOKEAN_BORDEN: {
        // P = tau F^{-T} (HyperelasticConstitutiveModel.h:82-87 over kirchhoff)
        const double J = mdet(F);
        M3 FinvT = mcofactor(F);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) FinvT.a[i][j] *= 1.0 / J;
        return mmul(bordenTau(mat, F, g), FinvT);
    }
    case ZMPM_MODEL_NEO_HOOKEAN: {
        // P = tau F^{-T} with tau from NeoHookean::kirchhoff (NeoHookean.cpp:23-48)
        double J = mdet(F);
        M3 FinvT = mcofactor(F);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) FinvT.a[i][j] *= 1.0 / J;
        double scale = mat.lambda * std::log(J) - mat.mu;
        M3 P;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) P.a[i][j] = mat.mu * F.a[i][j] + scale * FinvT.a[i][j];
        return P;
    }
    default:
        return mzero();
    }
}

// Strain energy density (for finite-difference pinning only).
double psiOf(const zmpm_material& mat, const M3& F, double g)
{
    switch (mat.model) {
    case ZMPM_MODEL_FIXED_COROTATED: { // CorotatedElasticity.cpp:65-70
        double J = mdet(F);
        M3 R = polarR(F);
        double s = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s += (F.a[i][j] - R.a[i][j]) * (F.a[i][j] - R.a[i][j]);
        return mat.mu * s + 0.5 * mat.lambda * (J - 1) * (J - 1);
    }
    case ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN: {
        double e = 0;
        qrStablePiola(mat, g, mat.kx, mat.ky, F, nullptr, &e);
        return e;
    }
    case ZMPM_MODEL_QR_ANISOTROPIC: {
        double e = 0;
        qrAnisoPiola(mat, g, mat.fiber_stiffness, F, nullptr, &e);
        return e;
    }
    case ZMPM_MODEL_NEO_HOOKEAN_BORDEN:
        return bordenPsi(mat, F, g);
    case ZMPM_MODEL_NEO_HOOKEAN: { // NeoHookean.cpp:31-35
        double J = mdet(F), lj = std::log(J), I1 = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) I1 += F.a[i][j] * F.a[i][j];
        return 0.5 * mat.mu * (I1 - 3) - mat.mu * lj + 0.5 * mat.lambda * lj * lj;
    }
    default:
        return 0;
    }
}

// V0 * kirchhoff stress: FBasedMpmForceHelper::updateState (FBasedMpmForceHelper.cpp:41-63)
// with kirchhoff = P F^T (HyperelasticConstitutiveModel.h:75-80); J-based:
// JBasedMpmForceHelper::updateState (JBasedMpmForceHelper.cpp:28-50).
M3 vtauOf(const zmpm_material& mat, const double* Fp, double V0, double g)
{
    M3 tau;
    if (mat.model == ZMPM_MODEL_EQUATION_OF_STATE) {
        double J = Fp[0];
        double dpsi_dJ = -mat.bulk * (std::pow(J, -mat.gamma) - 1.0); // EquationOfState.h:83-86
        double s = V0 * dpsi_dJ * J;
        tau = mzero();
        tau.a[0][0] = tau.a[1][1] = tau.a[2][2] = s;
        return tau;
    }
    M3 F = mload(Fp);
    if (mat.model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN)
        tau = bordenTau(mat, F, g); // the model's own kirchhoff (FBasedMpmForceHelper.cpp:58)
    else if (mat.model == ZMPM_MODEL_NEO_HOOKEAN) {
        // NeoHookean::kirchhoff (NeoHookean.cpp:38-43): tau = mu F F^T + (lambda logJ - mu) I
        double J = mdet(F);
        double scale = mat.lambda * std::log(J) - mat.mu;
        M3 FFt = mmul(F, mtrans(F));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) tau.a[i][j] = mat.mu * FFt.a[i][j] + (i == j ? scale : 0.0);
    }
    else {
        M3 P = firstPiola(mat, F, g);
        tau = mmul(P, mtrans(F));
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tau.a[i][j] *= V0;
    return tau;
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (columns of Vec
// are unit eigenvectors).  Stands in for Eigen::EigenSolver at
// AnisoFractureSimulation.h:244-254: sigma^+ is invariant to eigenvector sign and
// to the basis chosen inside a degenerate eigenspace.
void symEigen(const M3& S_in, double* lam, M3& Vec)
{
    M3 S = S_in;
    Vec = meye();
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = S.a[0][1] * S.a[0][1] + S.a[0][2] * S.a[0][2] + S.a[1][2] * S.a[1][2];
        double diag = S.a[0][0] * S.a[0][0] + S.a[1][1] * S.a[1][1] + S.a[2][2] * S.a[2][2];
        if (off <= 1e-36 * diag || off == 0) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                double apq = S.a[p][q];
                if (apq == 0) continue;
                double theta = (S.a[q][q] - S.a[p][p]) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; k++) { // S <- J^T S J
                    double skp = S.a[k][p], skq = S.a[k][q];
                    S.a[k][p] = c * skp - s * skq;
                    S.a[k][q] = s * skp + c * skq;
                }
                for (int k = 0; k < 3; k++) {
                    double spk = S.a[p][k], sqk = S.a[q][k];
                    S.a[p][k] = c * spk - s * sqk;
                    S.a[q][k] = s * spk + c * sqk;
                }
                for (int k = 0; k < 3; k++) {
                    double vkp = Vec.a[k][p], vkq = Vec.a[k][q];
                    Vec.a[k][p] = c * vkp - s * vkq;
                    Vec.a[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; i++) lam[i] = S.a[i][i];
}

inline double macaulay(double x) { return (x + std::fabs(x)) / 2; } // MathTools.h:276-281

// The crack driving force D~ = zeta <phi - 1>_+ of one particle
// (AnisoFractureSimulation.h:222-268, and :302-332 for the implicit solve) with
// constructStructuralTensor (:365-391) and evaluateEffectiveCauchyStress
// (QRStableNeoHookean.h:83-93, QRAnisotropic.h:79-91).
double dTildeOf(const zmpm_material& mat, const M3& F)
{
    // undegraded isotropic Cauchy stress: g = 1 and kx = ky = 0 (QR-NH) or every
    // fibre at k_fiber_isotropic = mu (QR-aniso)
    double J = 1;
    M3 cauchy;
    if (mat.model == ZMPM_MODEL_NEO_HOOKEAN_BORDEN) { // NeoHookeanBorden.cpp:96-108
        J = mdet(F);
        cauchy = bordenTau(mat, F, 1.0);
    }
    else {
        M3 P;
        if (mat.model == ZMPM_MODEL_QR_ANISOTROPIC) {
            const double kiso[3] = { mat.mu, mat.mu, mat.mu };
            P = qrAnisoPiola(mat, 1.0, kiso, F, &J, nullptr);
        }
        else
            P = qrStablePiola(mat, 1.0, 0.0, 0.0, F, &J, nullptr);
        cauchy = mmul(P, mtrans(F));
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cauchy.a[i][j] /= J;
    M3 sym;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) sym.a[i][j] = 0.5 * (cauchy.a[i][j] + cauchy.a[j][i]);
    double lam[3];
    M3 vec;
    symEigen(sym, lam, vec);
    M3 sp = mzero();
    for (int e = 0; e < 3; e++) {
        double l = macaulay(lam[e]);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) sp.a[i][j] += l * (vec.a[i][e] * vec.a[j][e]);
    }
    // structural tensor A = I + sum_k alpha_k (R a0_k)(R a0_k)^T, R from the SVD of F
    M3 R = polarR(F);
    M3 A = meye();
    int nd = mat.num_directors == 2 ? 2 : 1;
    for (int k = 0; k < nd; k++) {
        double a1[3];
        for (int i = 0; i < 3; i++) a1[i] = R.a[i][0] * mat.a0[k][0] + R.a[i][1] * mat.a0[k][1] + R.a[i][2] * mat.a0[k][2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A.a[i][j] += mat.alphas[k] * (a1[i] * a1[j]);
    }
    M3 Asig = mmul(A, sp), sigA = mmul(sp, A);
    double contraction = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) contraction += Asig.a[i][j] * sigA.a[i][j];
    double phi = (1.0 / (mat.sigma_crit * mat.sigma_crit)) * contraction;
    return mat.zeta * macaulay(phi - 1);
}

// updateDamageHelper for one particle (AnisoFractureSimulation.h:210-283).  dt
// is the substep.
void damageUpdate(const zmpm_material& mat, const M3& F, double lap, double dt, double& d, double& g)
{
    double geometricResist = d - (mat.l0 * mat.l0 * lap);
    double dTilde = dTildeOf(mat, F);
    double diff = ((1 - d) * dTilde) - geometricResist;
    double newD = d + ((dt / mat.eta) * macaulay(diff));
    d = std::min(newD, 1.0);
    double k = mat.residual_phase;
    g = ((1 - d) * (1 - d) * (1 - k)) + k;
}

// ------------------------------------------------------------ SPGrid layout
// SPGrid_Mask_base / SPGrid_Mask (Lib/SPGrid/Core/SPGrid_Mask.h:22-129) for
// log2_page = 12 and the GridState<T,3> struct size (MpmGrid.h:16-36):
// float -> 64 B (log2 6), double -> 128 B (log2 7).
struct Layout {
    int data_bits, block_bits, xbits, ybits, zbits;
    uint64_t xmask, ymask, zmask;
};

Layout makeLayout(int log2_struct)
{
    Layout L;
    const int page_bits = 12;
    L.data_bits = log2_struct;
    L.block_bits = page_bits - L.data_bits;
    L.zbits = L.block_bits / 3 + (L.block_bits % 3 > 0);
    L.ybits = L.block_bits / 3 + (L.block_bits % 3 > 1);
    L.xbits = L.block_bits / 3;
    const int sh = 3 - L.block_bits % 3 + (page_bits - 12);
    const uint64_t hi = ~0ull << page_bits;
    uint64_t page_z = (0x9249249249249249ull << sh) & hi;
    uint64_t page_y = (0x2492492492492492ull << sh) & hi;
    uint64_t page_x = (0x4924924924924924ull << sh) & hi;
    const int log2_field = log2_struct; // Array_type<> covers the whole struct
    uint64_t ez = (uint64_t)((1 << L.zbits) - 1) << log2_field;
    uint64_t ey = (uint64_t)((1 << L.ybits) - 1) << (log2_field + L.zbits);
    uint64_t ex = (uint64_t)((1 << L.xbits) - 1) << (log2_field + L.zbits + L.ybits);
    L.zmask = page_z | ez;
    L.ymask = page_y | ey;
    L.xmask = page_x | ex;
    return L;
}

// Bit_Spread_u64 (SPGrid_Utilities.h:81-131): deposit the low bits of `data`
// (sign-extended to 64 bits) into the set bits of `mask`, low to high.
uint64_t bitSpread(int64_t data, uint64_t mask)
{
    uint64_t u = (uint64_t)data, result = 0;
    for (int b = 0; b < 64; b++) {
        uint64_t bit = 1ull << b;
        if (mask & bit)
            result |= u & bit;
        else
            u <<= 1;
    }
    return result;
}

// Bit_Pack (SPGrid_Utilities.cpp:43-62): gather the bits of `data` under `mask`.
int64_t bitPack(uint64_t data, uint64_t mask)
{
    int64_t r = 0;
    int out = 0;
    for (int b = 0; b < 64; b++)
        if (mask & (1ull << b)) {
            if (data & (1ull << b)) r |= (int64_t)1 << out;
            out++;
        }
    return r;
}

inline uint64_t linearOffset(const Layout& L, int i, int j, int k)
{
    return bitSpread(i, L.xmask) | bitSpread(j, L.ymask) | bitSpread(k, L.zmask);
}

// Packed_Add (SPGrid_Mask.h:237-245)
inline uint64_t packedAdd(const Layout& L, uint64_t a, uint64_t b)
{
    uint64_t MX = ~L.xmask, MY = ~L.ymask, MZ = ~L.zmask, W = L.xmask | L.ymask | L.zmask;
    uint64_t x = ((a | MX) + (b & ~MX)) & ~MX;
    uint64_t y = ((a | MY) + (b & ~MY)) & ~MY;
    uint64_t z = ((a | MZ) + (b & ~MZ)) & ~MZ;
    uint64_t w = ((a | W) + (b & ~W)) & ~W;
    return x | y | z | w;
}

// int_floor (Math/MathTools.h:15-25)
inline int intFloor(double x)
{
    int i = (int)x;
    return i - (i > x);
}

// Quadratic B-spline (BSplines.h:72-98): base node, weights, gradients, Hessian diag.
inline void bspline(double x, int& base, double* w, double* dw, double* ddw)
{
    base = intFloor(x - 0.5);
    double d0 = x - base;
    double z = 1.5 - d0;
    double z2 = z * z;
    w[0] = 0.5 * z2;
    double d1 = d0 - 1;
    w[1] = 0.75 - d1 * d1;
    double d2 = 1 - d1;
    double zz = 1.5 - d2;
    double zz2 = zz * zz;
    w[2] = 0.5 * zz2;
    if (dw) {
        dw[0] = -z;
        dw[1] = -2 * d1;
        dw[2] = zz;
    }
    if (ddw) {
        ddw[0] = 1;
        ddw[1] = -2;
        ddw[2] = 1;
    }
}

// ------------------------------------------------------------------- grid
struct Node {
    double v[3];
    double m;
    double new_v[3];
    int64_t idx;
    double pf, pfm;
};

struct Collider {
    zmpm_collider c;
    double n[3]; // normalised half-space normal
};

// ------------------------------------------------------- collision objects
// Eigen's Quaternion::toRotationMatrix for q = (w, x, y, z); an all-zero q
// stands for the identity.
M3 quatToMatrix(const double* q)
{
    double w = q[0], x = q[1], y = q[2], z = q[3];
    if (w == 0 && x == 0 && y == 0 && z == 0) w = 1;
    const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x;
    const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
    M3 R;
    R.a[0][0] = 1 - (tyy + tzz);
    R.a[0][1] = txy - twz;
    R.a[0][2] = txz + twy;
    R.a[1][0] = txy + twz;
    R.a[1][1] = 1 - (txx + tzz);
    R.a[1][2] = tyz - twx;
    R.a[2][0] = txz - twy;
    R.a[2][1] = tyz + twx;
    R.a[2][2] = 1 - (txx + tyy);
    return R;
}

// State of a collision object at time t, as the updateState callbacks of the
// examples set it (examples/boneTwist.h:144-170): b = b0 + dbdt t, a rotation
// q(t) = (cos(|w| t / 2), sin(|w| t / 2) w / |w|) * q0 about a constant angular
// velocity w, s = s0 + dsdt t (AnalyticCollisionObject::setTranslation /
// setRotation / setAngularVelocity / setScaling, CollisionObject.h:144-166).
struct ColliderAt {
    M3 R;
    double b[3], s;
};

ColliderAt colliderAt(const zmpm_collider& c, double t)
{
    ColliderAt st;
    for (int a = 0; a < 3; a++) st.b[a] = c.b0[a] + c.dbdt[a] * t;
    st.s = (c.scale0 == 0 ? 1.0 : c.scale0) + c.dsdt * t;
    double q0[4] = { c.rotation[0], c.rotation[1], c.rotation[2], c.rotation[3] };
    if (q0[0] == 0 && q0[1] == 0 && q0[2] == 0 && q0[3] == 0) q0[0] = 1;
    const double wn = std::sqrt(c.omega[0] * c.omega[0] + c.omega[1] * c.omega[1] + c.omega[2] * c.omega[2]);
    double q[4] = { q0[0], q0[1], q0[2], q0[3] };
    if (wn > 0) {
        const double ha = 0.5 * wn * t, cw = std::cos(ha), sw = std::sin(ha) / wn;
        const double p[4] = { cw, sw * c.omega[0], sw * c.omega[1], sw * c.omega[2] };
        // Hamilton product p * q0
        q[0] = p[0] * q0[0] - p[1] * q0[1] - p[2] * q0[2] - p[3] * q0[3];
        q[1] = p[0] * q0[1] + p[1] * q0[0] + p[2] * q0[3] - p[3] * q0[2];
        q[2] = p[0] * q0[2] - p[1] * q0[3] + p[2] * q0[0] + p[3] * q0[1];
        q[3] = p[0] * q0[3] + p[1] * q0[2] - p[2] * q0[1] + p[3] * q0[0];
    }
    st.R = quatToMatrix(q);
    return st;
}

// AnalyticLevelSet::queryInside in material space X: inside flag, signed
// distance and the material-space normal N.
//  half space  HalfSpace::signedDistance / normal (AnalyticLevelSet.cpp:331-347)
//  box         AxisAlignedAnalyticBox -> AnalyticBox::signedDistancePrimitive and its
//              AutoDiff normal (AnalyticLevelSet.cpp:412-428, 562-588): d = |X - c| - h;
//              inside (max d <= 0) the normal is sign(X_a - c_a) e_a of the first
//              largest d_a (Eigen's maxCoeff keeps the first; std::min(dd, 0) keeps dd
//              at dd = 0; AutoDiff abs takes sign +1 at 0, AutoDiff.h:307-311, and the
//              norm of the clamped d has derivative 0 at 0, AutoDiff.h:319-324);
//              outside it is sum_a max(d_a, 0) sign(X_a - c_a) e_a / |max(d, 0)|
//  sphere      Sphere::queryInside (AnalyticLevelSet.cpp:494-511): inside when
//              |X - c|^2 < r^2, N = (X - c)/|X - c|, e_x when |X - c| < 1e-7
//  cylinder    CappedCylinder::signedDistancePrimitive and its AutoDiff normal
//              (AnalyticLevelSet.h:290-320) in the primitive frame R_c^-1 (X - b_c):
//              q = (|(x, z)|, y), d = |q| - (r, h/2); inside the normal is the
//              gradient of the larger d (d_0 on a tie), outside that of |max(d, 0)|
bool levelSetInside(const Collider& co, const double X[3], double& phi, double N[3])
{
    const zmpm_collider& c = co.c;
    N[0] = N[1] = N[2] = 0;
    if (c.shape == ZMPM_SHAPE_HALF_SPACE) {
        phi = co.n[0] * (X[0] - c.origin[0]) + co.n[1] * (X[1] - c.origin[1]) + co.n[2] * (X[2] - c.origin[2]);
        for (int a = 0; a < 3; a++) N[a] = co.n[a];
        return phi <= 0;
    }
    if (c.shape == ZMPM_SHAPE_BOX) {
        double dd[3], Xp[3];
        for (int a = 0; a < 3; a++) {
            const double cen = (c.box_min[a] + c.box_max[a]) / 2;
            const double he = (c.box_max[a] - c.box_min[a]) / 2;
            Xp[a] = X[a] - cen;
            dd[a] = std::fabs(Xp[a]) - he;
        }
        int am = 0;
        for (int a = 1; a < 3; a++)
            if (dd[a] > dd[am]) am = a;
        double q2 = 0;
        for (int a = 0; a < 3; a++) q2 += dd[a] > 0 ? dd[a] * dd[a] : 0.0;
        phi = std::min(dd[am], 0.0) + std::sqrt(q2);
        if (q2 > 0) {
            const double qn = std::sqrt(q2);
            for (int a = 0; a < 3; a++) N[a] = (dd[a] > 0 ? dd[a] / qn : 0.0) * (Xp[a] < 0 ? -1.0 : 1.0);
        }
        else
            N[am] = Xp[am] < 0 ? -1.0 : 1.0;
        return phi <= 0;
    }
    if (c.shape == ZMPM_SHAPE_SPHERE) {
        double tc[3], d2 = 0;
        for (int a = 0; a < 3; a++) {
            tc[a] = X[a] - c.origin[a];
            d2 += tc[a] * tc[a];
        }
        const double dist = std::sqrt(d2);
        phi = dist - c.radius;
        if (dist < 1e-7)
            N[0] = 1;
        else
            for (int a = 0; a < 3; a++) N[a] = (1 / dist) * tc[a];
        return d2 < c.radius * c.radius;
    }
    // capped cylinder
    const M3 Rc = quatToMatrix(c.shape_rotation);
    const M3 Ri = minverse(Rc);
    double Y[3], Xp[3];
    for (int a = 0; a < 3; a++) Y[a] = X[a] - c.origin[a];
    for (int a = 0; a < 3; a++) Xp[a] = Ri.a[a][0] * Y[0] + Ri.a[a][1] * Y[1] + Ri.a[a][2] * Y[2];
    const double rxz = std::sqrt(Xp[0] * Xp[0] + Xp[2] * Xp[2]);
    const double d0 = rxz - c.radius, d1 = std::fabs(Xp[1]) - 0.5 * c.height;
    // gradients in the primitive frame
    double g0[3] = { 0, 0, 0 }, g1[3] = { 0, Xp[1] < 0 ? -1.0 : 1.0, 0 };
    if (rxz != 0) {
        g0[0] = Xp[0] / rxz;
        g0[2] = Xp[2] / rxz;
    }
    const double q0 = std::max(d0, 0.0), q1 = std::max(d1, 0.0);
    const double mx = std::max(d0, d1);
    const double qn = std::sqrt(q0 * q0 + q1 * q1);
    phi = std::min(mx, 0.0) + qn;
    double n[3] = { 0, 0, 0 };
    if (!(0 < mx)) { // min(max(d0, d1), 0) = max(d0, d1), whose gradient is d0's on a tie
        const double* g = (d0 < d1) ? g1 : g0;
        for (int a = 0; a < 3; a++) n[a] = g[a];
    }
    if (qn != 0)
        for (int a = 0; a < 3; a++) n[a] += ((d0 < 0 ? 0.0 : d0) * g0[a] + (d1 < 0 ? 0.0 : d1) * g1[a]) / qn;
    for (int a = 0; a < 3; a++) N[a] = Rc.a[a][0] * n[0] + Rc.a[a][1] * n[1] + Rc.a[a][2] * n[2];
    return phi <= 0;
}

// AnalyticCollisionObject::detectAndResolveCollision (CollisionObject.cpp:341-406)
// for the object state st: true when x is inside; vi resolved; n the world normal
// of a SLIP / SEPARATE object.
bool detectAndResolve(const Collider& co, const ColliderAt& st, const double* xi, double* vi, double* n)
{
    const zmpm_collider& c = co.c;
    double xb[3], X[3];
    for (int a = 0; a < 3; a++) xb[a] = xi[a] - st.b[a];
    const double os = 1 / st.s;
    for (int a = 0; a < 3; a++) X[a] = (st.R.a[0][a] * xb[0] + st.R.a[1][a] * xb[1] + st.R.a[2][a] * xb[2]) * os;
    double phi, N[3];
    if (!levelSetInside(co, X, phi, N)) return false;
    // v_object = omega x (x - b) + (dsdt / s)(x - b) + dbdt
    double vo[3];
    const double* w = c.omega;
    const double cr[3] = { w[1] * xb[2] - w[2] * xb[1], w[2] * xb[0] - w[0] * xb[2], w[0] * xb[1] - w[1] * xb[0] };
    for (int a = 0; a < 3; a++) vo[a] = cr[a] + (c.dsdt * os) * xb[a] + c.dbdt[a];
    for (int a = 0; a < 3; a++) vi[a] -= vo[a];
    if (c.type == ZMPM_STICKY)
        vi[0] = vi[1] = vi[2] = 0;
    else {
        for (int a = 0; a < 3; a++) n[a] = st.R.a[a][0] * N[0] + st.R.a[a][1] * N[1] + st.R.a[a][2] * N[2];
        const double dot = vi[0] * n[0] + vi[1] * n[1] + vi[2] * n[2];
        if (c.type == ZMPM_SLIP || dot < 0) {
            for (int a = 0; a < 3; a++) vi[a] -= n[a] * dot;
            // kinematic friction, only against an approaching velocity
            if (c.friction != 0 && dot < 0) {
                const double vn = std::sqrt(vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
                if (-dot * c.friction < vn)
                    for (int a = 0; a < 3; a++) vi[a] += (vi[a] / vn) * dot * c.friction;
                else
                    vi[0] = vi[1] = vi[2] = 0;
            }
        }
    }
    for (int a = 0; a < 3; a++) vi[a] += vo[a];
    return true;
}

// AnalyticCollisionObject::multiObjectCollision (CollisionObject.cpp:66-105) at
// time t: objects in insertion order; the first STICKY hit ends the loop; the
// normals of SLIP / SEPARATE hits are Gram-Schmidt orthonormalised into a basis
// and the loop ends once it holds 3 of them (a hit whose normal lies in the span
// of the earlier ones adds none).
// The objects' states at the substep's time (the reference's updateState at
// reinitialize, MpmSimulationBase.cpp:209-213).
std::vector<ColliderAt> collidersAt(const std::vector<Collider>& cols, double t)
{
    std::vector<ColliderAt> st;
    for (const Collider& co : cols) st.push_back(colliderAt(co.c, t));
    return st;
}

bool multiObjectCollision(const std::vector<Collider>& cols, const std::vector<ColliderAt>& states, const double* xi,
                          double* vi)
{
    bool any = false;
    int count = 0;
    double basis[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
    for (size_t k = 0; k < cols.size(); k++) {
        const Collider& co = cols[k];
        if (co.c.type == ZMPM_GHOST) continue;
        double n[3] = { 0, 0, 0 };
        if (!detectAndResolve(co, states[k], xi, vi, n)) continue;
        any = true;
        if (co.c.type == ZMPM_STICKY) break;
        for (int k = 0; k < count; k++) {
            const double dot = basis[k][0] * n[0] + basis[k][1] * n[1] + basis[k][2] * n[2];
            for (int a = 0; a < 3; a++) n[a] -= dot * basis[k][a];
        }
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len != 0) {
            for (int a = 0; a < 3; a++) basis[count][a] = n[a] / len;
            if (++count == 3) break;
        }
    }
    return any;
}

} // namespace

// --------------------------------------------------------------- the handle
struct zmpm_context {
    zmpm_scene_desc desc;
    Layout L;
    std::string err;
    zmpm_launch_config launch { 1, 1, 1, 1, 0, { 0, 0, 0 } };
    std::vector<zmpm_material> mats;
    std::vector<Collider> colliders;
    bool any_damage = false;
    // particle state (index order, never reordered)
    int64_t n = 0;
    std::vector<double> x, v, C, F, m, V0, d, g, lap, ys, dtil;
    // FJ-mixed particles: the "J" attribute and the J model's lambda (per
    // particle, as the reference keeps one QuadraticVolumePenalty per particle)
    std::vector<double> J, jl;
    int fj_state = 0; // FJMixedMpmForceHelper::state: 0 SOLVE_F, 1 SOLVE_J
    // MpmSimulationBase::fext as addExternalBodyForce builds its entries
    // (scale, frequency, direction), and explicit_velocity_field as v = a + B x
    struct BodyForce {
        double scale, freq, dir[3];
    };
    std::vector<BodyForce> fext;
    bool vfield = false;
    double va[3] = { 0, 0, 0 }, vB[9] = {};
    // zmpm_stage: the next stage of the substep in progress (0 = none), its dt and t
    int stage_next = 0;
    double stage_dt = 0, stage_t = 0;
    std::vector<int32_t> mat;
    int cg_iterations = -1; // last implicit damage solve
    // binning of the last substep
    std::vector<uint64_t> keys;
    std::vector<int32_t> order;
    std::vector<std::pair<int, int>> groups;
    std::vector<uint64_t> block_offset;
    std::vector<uint64_t> base_offset;
    std::vector<int32_t> base_node; // 3 per particle: base node of the substep's sort
    // grid (page map in first-touch order, MpmGrid::pollute)
    std::unordered_map<uint64_t, int> page_to_block;
    std::vector<uint64_t> block_pages;
    std::vector<Node> nodes;
    int NB = 0;
    int latched = ZMPM_OK;
    bool stepped = false;
    // a batch (zmpm_create_batch): one single-scene context per scene, run one
    // after the other; particles keep their upload order through `pscene`
    std::vector<zmpm_context*> subs;
    std::vector<std::array<int, 3>> origins;
    std::vector<int32_t> pscene;
    ~zmpm_context();
};
zmpm_context::~zmpm_context()
{
    for (zmpm_context* c : subs) delete c;
}

namespace {

zmpm_status fail(zmpm_context* h, zmpm_status s, const std::string& msg)
{
    if (h) h->err = msg;
    return s;
}

// FP32 layout position grid: x / dx in units of 2^-pb, pb = 21 for domains of
// up to 1024 nodes per axis and 19 up to 4096, clamped to [-2^30, 2^31 - 2]
// units (outside the domain either way).
inline int posBitsOf(int domain_nodes) { return domain_nodes <= 1024 ? 21 : 19; }

inline double posUnits(double x, double dx, int pb)
{
    const double u = std::nearbyint(x / dx * (double)(1 << pb));
    return std::min(std::max(u, -1073741824.0), 2147483646.0);
}

inline double quantizePos(double x, double dx, int pb)
{
    if (!std::isfinite(x)) return x;
    return posUnits(x, dx, pb) * dx / (double)(1 << pb);
}

// Base node of particle i in the handle's precision arithmetic; false if the
// node leaves the addressable domain.
bool baseNodeOf(const zmpm_context* h, int64_t i, int* b, bool* nan_out)
{
    bool ok = true;
    for (int a = 0; a < 3; a++) {
        double xv = h->x[3 * i + a];
        if (!std::isfinite(xv)) {
            *nan_out = true;
            b[a] = 0;
            ok = false;
            continue;
        }
        double s;
        if (h->desc.precision == ZMPM_PREC_FP32) {
            // floor(X / 2^pb - 1/2) of the integer X, exactly
            const int pb = posBitsOf(h->desc.domain_nodes);
            const int64_t X = (int64_t)posUnits(xv, h->desc.dx, pb);
            const int64_t t = X - (int64_t(1) << (pb - 1));
            b[a] = (int)((t >= 0) ? (t >> pb) : -((-t + (int64_t(1) << pb) - 1) >> pb));
        }
        else {
            double odx = 1 / h->desc.dx;
            s = xv * odx;
            double t = s - 0.5;
            if (!(t > -1e9 && t < 1e9)) { ok = false; b[a] = 0; continue; }
            b[a] = intFloor(t);
        }
        if (b[a] < 0 || b[a] > h->desc.domain_nodes - 5) ok = false;
    }
    return ok;
}

// Index-space position in the handle's precision (BSplineWeights::compute, MpmGrid.h:74-79).
inline double indexSpace(const zmpm_context* h, double xv)
{
    if (h->desc.precision == ZMPM_PREC_FP32) {
        const int pb = posBitsOf(h->desc.domain_nodes);
        return posUnits(xv, h->desc.dx, pb) / (double)(1 << pb);
    }
    return xv * (1 / h->desc.dx);
}

inline Node* nodeAt(zmpm_context* h, int i, int j, int k)
{
    uint64_t off = linearOffset(h->L, i, j, k);
    auto it = h->page_to_block.find(off >> 12);
    if (it == h->page_to_block.end()) return nullptr;
    int elem = (int)((off >> h->L.data_bits) & (uint64_t)(h->NB - 1));
    return &h->nodes[(size_t)it->second * h->NB + elem];
}

void setPage(zmpm_context* h, uint64_t offset)
{
    uint64_t page = offset >> 12;
    if (h->page_to_block.find(page) == h->page_to_block.end()) {
        h->page_to_block.emplace(page, (int)h->block_pages.size());
        h->block_pages.push_back(page);
    }
}

// sortParticlesAndPolluteGrid (MpmSimulationBase.cpp:1531-1573) + MpmGrid::pollute (MpmGrid.h:153-193)
zmpm_status sortAndPollute(zmpm_context* h)
{
    const Layout& L = h->L;
    const int64_t n = h->n;
    const int index_bits = 32 - L.block_bits;
    if (n >= ((int64_t)1 << index_bits)) return fail(h, ZMPM_E_ARG, "too many particles for the key layout");
    h->keys.resize(n);
    h->base_offset.resize(n);
    h->base_node.resize(3 * n);
    h->order.resize(n);
    bool bad = false, nan = false;
    for (int64_t i = 0; i < n; i++) {
        int b[3];
        bool isnan = false;
        if (!baseNodeOf(h, i, b, &isnan)) {
            bad = true;
            nan = nan || isnan;
            continue;
        }
        uint64_t off = linearOffset(L, b[0], b[1], b[2]);
        h->base_offset[i] = off;
        for (int a = 0; a < 3; a++) h->base_node[3 * i + a] = b[a];
        h->keys[i] = ((off >> L.data_bits) << index_bits) + (uint64_t)i;
    }
    if (bad) {
        h->latched = nan ? ZMPM_E_NAN : ZMPM_E_DOMAIN;
        return fail(h, (zmpm_status)h->latched, nan ? "non-finite particle position" : "particle outside the addressable domain");
    }
    std::sort(h->keys.begin(), h->keys.end());
    h->groups.clear();
    h->block_offset.clear();
    int last = 0;
    for (int64_t i = 0; i < n; i++) {
        if (i == n - 1 || (h->keys[i] >> 32) != (h->keys[i + 1] >> 32)) {
            h->groups.emplace_back(last, (int)i);
            h->block_offset.push_back(h->keys[i] >> 32);
            last = (int)i + 1;
        }
        h->order[i] = (int32_t)(h->keys[i] & ((1ull << index_bits) - 1));
    }
    // pollute: first-touch page order over particles in index order
    h->page_to_block.clear();
    h->block_pages.clear();
    const uint64_t bx = 1u << L.xbits, by = 1u << L.ybits, bz = 1u << L.zbits;
    for (int64_t i = 0; i < n; i++) {
        uint64_t off = h->base_offset[i];
        if (!i || off != h->base_offset[i - 1]) {
            setPage(h, off);
            for (int a = 0; a <= 1; a++)
                for (int b = 0; b <= 1; b++)
                    for (int c = 0; c <= 1; c++)
                        setPage(h, packedAdd(L, off, linearOffset(L, (int)(bx * a), (int)(by * b), (int)(bz * c))));
        }
    }
    h->nodes.assign(h->block_pages.size() * (size_t)h->NB, Node());
    for (auto& nd : h->nodes) {
        std::memset(&nd, 0, sizeof(Node));
        nd.idx = -1;
    }
    return ZMPM_OK;
}

// Weights of the 27-node stencil in MpmGrid::iterateKernel order (MpmGrid.h:437-488).
struct Stencil {
    int base[3];
    double w[3][3], dw[3][3], ddw[3][3];
    double odx;
};

inline void makeStencil(const zmpm_context* h, const double* xp, Stencil& s)
{
    s.odx = 1 / h->desc.dx;
    for (int a = 0; a < 3; a++) bspline(indexSpace(h, xp[a]), s.base[a], s.w[a], s.dw[a], s.ddw[a]);
}

// addr: the base node whose stencil the grid memory is taken from; null for
// s.base.  The node coordinates passed to op are always s.base + (i, j, k):
// iterateKernel reads memory at particle_base_offset but reports the spline's
// own base node (MpmGrid.h:437-488).
template <class OP>
inline void iterateKernel(zmpm_context* h, const Stencil& s, const OP& op, const int* addr = nullptr)
{
    const int* ab = addr ? addr : s.base;
    const double odx = s.odx;
    for (int i = 0; i < 3; i++) {
        double wi = s.w[0][i];
        double dwidxi = odx * s.dw[0][i];
        for (int j = 0; j < 3; j++) {
            double wj = s.w[1][j];
            double wij = wi * wj;
            double dwijdxi = dwidxi * wj;
            double dwijdxj = wi * odx * s.dw[1][j];
            for (int k = 0; k < 3; k++) {
                double wk = s.w[2][k];
                double wijk = wij * wk;
                double dw3[3] = { dwijdxi * wk, dwijdxj * wk, wij * odx * s.dw[2][k] };
                // Laplacian weight with the 3D index quirk (MpmGrid.h:608-613):
                // ddw[1] and ddw[2] are indexed by i, not j and k.
                double t1 = odx * odx * s.ddw[0][i] * wj * wk;
                double t2 = odx * odx * s.ddw[1][i] * wi * wk;
                double t3 = odx * odx * s.ddw[2][i] * wi * wj;
                int node[3] = { s.base[0] + i, s.base[1] + j, s.base[2] + k };
                Node* g = nodeAt(h, ab[0] + i, ab[1] + j, ab[2] + k);
                op(node, wijk, dw3, t1 + t2 + t3, *g);
            }
        }
    }
}

const zmpm_material& matOf(const zmpm_context* h, int64_t i) { return h->mats[(size_t)h->mat[i]]; }

// VonMisesStvkHencky::projectStrain (PlasticityApplier.cpp:1480-1516): SVD of
// F, Hencky strain eps = log(max(sigma, 1e-4)), its deviator eps_hat; when
// |eps_hat| exceeds yield / (2 mu) the strain returns radially onto it,
// F = U exp(H) V^T, and the yield stress hardens by xi delta_gamma.  The
// fail-stress bookkeeping (a crack normal the explicit step never reads) is
// not kept.
void vonMisesProject(double mu, double xi, double& yield_stress, M3& F)
{
    M3 U, V;
    double sigma[3];
    svd3(F, U, sigma, V, kSvdTol);
    double eps[3], tr = 0;
    for (int a = 0; a < 3; a++) {
        eps[a] = std::log(std::max(sigma[a], 1e-4));
        tr += eps[a];
    }
    double hat[3], n2 = 0;
    for (int a = 0; a < 3; a++) {
        hat[a] = eps[a] - (tr / 3.0);
        n2 += hat[a] * hat[a];
    }
    const double nrm = std::sqrt(n2);
    const double delta_gamma = nrm - yield_stress / (2 * mu);
    if (delta_gamma <= 0) return;
    yield_stress -= xi * delta_gamma;
    double e[3];
    for (int a = 0; a < 3; a++) e[a] = std::exp(eps[a] - (delta_gamma / nrm) * hat[a]);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F.a[r][c] = U.a[r][0] * e[0] * V.a[c][0] + U.a[r][1] * e[1] * V.a[c][1] + U.a[r][2] * e[2] * V.a[c][2];
}

// DruckerPragerStvkHencky (PlasticityApplier.cpp:811-876): alpha from the
// friction angle as its constructor computes it; projectStrain on the Hencky
// strain eps = log(max(|sigma|, 1e-4)) - cohesion with the per-particle log Jp:
// tip (tr >= 0: F = exp(cohesion) U V^T, log Jp += beta tr eps when volume
// correction is on), inside (H = eps + cohesion) or onto the cone.
double dpAlpha(double friction_angle)
{
    const double sin_phi = std::sin(friction_angle / 180.0 * 3.141592653);
    return std::sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi);
}

void druckerPragerProject(const zmpm_material& mt, double& logJp, M3& F)
{
    M3 U, V;
    double sigma[3];
    svd3(F, U, sigma, V, kSvdTol);
    const double coh = mt.cohesion;
    double eps[3], sum = 0;
    for (int a = 0; a < 3; a++) {
        eps[a] = std::log(std::max(std::fabs(sigma[a]), 1e-4)) - coh;
        sum += eps[a];
    }
    const double tr = sum + logJp;
    double hat[3], n2 = 0;
    for (int a = 0; a < 3; a++) {
        hat[a] = eps[a] - (tr / 3.0);
        n2 += hat[a] * hat[a];
    }
    double e[3];
    if (tr >= 0) {
        for (int a = 0; a < 3; a++) e[a] = std::exp(coh);
        if (mt.volume_correction) logJp = mt.hardening * sum + logJp;
    }
    else if (mt.mu == 0)
        return;
    else {
        logJp = 0;
        const double nrm = std::sqrt(n2);
        const double dgam = nrm + (3 * mt.lambda + 2 * mt.mu) / (2 * mt.mu) * tr * dpAlpha(mt.friction_angle);
        for (int a = 0; a < 3; a++) e[a] = std::exp(dgam <= 0 ? eps[a] + coh : eps[a] - (dgam / nrm) * hat[a] + coh);
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F.a[r][c] = U.a[r][0] * e[0] * V.a[c][0] + U.a[r][1] * e[1] * V.a[c][1] + U.a[r][2] * e[2] * V.a[c][2];
}

// SnowPlasticity::projectStrain (PlasticityApplier.cpp:16-50): singular values
// clamped to [1 - theta_c, 1 + theta_s], F = U Sigma_clamped V^T, Jp <- Jp det F /
// det Fe clamped to [min_Jp, max_Jp] (the !(<=) / !(>=) tests of the reference,
// so a NaN lands on a bound).  The reference then scales the particle's mu and
// lambda by exp(psi (Jp_old - Jp_new)); over a run that is exp(psi (1 - Jp)) with
// Jp starting at 1, which snowScale gives from the stored Jp.
void snowProject(const zmpm_material& mt, double& Jp, M3& F)
{
    M3 U, V;
    double sigma[3];
    svd3(F, U, sigma, V, kSvdTol);
    double Fe_det = 1;
    for (int a = 0; a < 3; a++) {
        sigma[a] = std::max(std::min(sigma[a], 1.0 + mt.critical_stretch), 1.0 - mt.critical_compression);
        Fe_det *= sigma[a];
    }
    double Jp_new = Jp * mdet(F) / Fe_det;
    if (!(Jp_new <= mt.max_jp)) Jp_new = mt.max_jp;
    if (!(Jp_new >= mt.min_jp)) Jp_new = mt.min_jp;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            F.a[r][c] = U.a[r][0] * sigma[0] * V.a[c][0] + U.a[r][1] * sigma[1] * V.a[c][1] + U.a[r][2] * sigma[2] * V.a[c][2];
    Jp = Jp_new;
}

// The factor on mu and lambda of a snow particle (1 for every other material).
double snowScale(const zmpm_material& mt, double Jp)
{
    return mt.plasticity == ZMPM_PLASTICITY_SNOW ? std::exp(mt.hardening * (1.0 - Jp)) : 1.0;
}

// gridToParticlesHelper<scheme, deg 2, mls> generic path (MpmSimulationBase.cpp:1412-1478)
// and evolveStrain: F <- (I + dt gradV) F (FBasedMpmForceHelper.cpp:93-108);
// J <- (1 + dt tr gradV) J (JBasedMpmForceHelper.cpp:79-94).  APIC_blend_RPIC:
// v = sum w v_i, C from B = sum w v_i (x_i - x_p)^T, grad v = sum v_i grad(w)^T
// (MLS: grad v = C).  FLIP_blend_PIC: v <- ratio v + sum w v_i - ratio sum w v_i^old,
// grad v = sum v_i grad(w)^T; C, which the reference does not keep for FLIP, is
// stored as 0.  old_base: grid memory through the substep's base offsets (the
// damping passes, whose particles have moved).  Returns the CFL flag.
int gridToParticlesPass(zmpm_context* h, double dt, double r, double flip, bool old_base)
{
    const int64_t n = h->n;
    const double dx = h->desc.dx;
    const double dx2 = dx * dx;
    const double D_inverse = 4 / dx2;
    const bool apic = h->desc.transfer_scheme == ZMPM_APIC_BLEND_RPIC;
    const bool mls = h->desc.mls_mpm != 0;
    const bool fp32_layout = h->desc.precision == ZMPM_PREC_FP32;
    int cfl_flag = 0;
    std::vector<double> gradV((size_t)n * 9);
#pragma omp parallel for schedule(static) reduction(| : cfl_flag)
    for (int64_t i = 0; i < n; i++) {
        double* Xp = &h->x[3 * i];
        double picV[3] = { 0, 0, 0 }, oldV[3] = { 0, 0, 0 }, Bp[3][3] = {}, gv[3][3] = {};
        Stencil s;
        makeStencil(h, Xp, s);
        iterateKernel(h, s, [&](const int* node, double w, const double* dw, double, Node& g) {
            double xi_xp[3];
            for (int a = 0; a < 3; a++) xi_xp[a] = (double)node[a] * dx - Xp[a];
            for (int a = 0; a < 3; a++) picV[a] += w * g.new_v[a];
            if (!apic)
                for (int a = 0; a < 3; a++) oldV[a] += w * g.v[a];
            for (int a = 0; a < 3; a++) {
                double wv = w * g.new_v[a];
                for (int c = 0; c < 3; c++) {
                    Bp[a][c] += wv * xi_xp[c];
                    gv[a][c] += g.new_v[a] * dw[c];
                }
            }
        }, old_base ? &h->base_node[3 * i] : nullptr);
        if (apic) {
            for (int a = 0; a < 3; a++) h->v[3 * i + a] = picV[a];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) {
                    double CC = Bp[a][c] * D_inverse, CCt = Bp[c][a] * D_inverse;
                    h->C[9 * i + 3 * a + c] = ((r + 1) * 0.5) * CC + ((r - 1) * 0.5) * CCt;
                    gradV[9 * i + 3 * a + c] = mls ? h->C[9 * i + 3 * a + c] : gv[a][c];
                }
        }
        else {
            for (int a = 0; a < 3; a++) {
                double& va = h->v[3 * i + a];
                va *= flip;
                va += picV[a] - flip * oldV[a];
            }
            for (int k = 0; k < 9; k++) {
                h->C[9 * i + k] = 0.0;
                gradV[9 * i + k] = gv[k / 3][k % 3];
            }
        }
        double inc[3] = { dt * picV[0], dt * picV[1], dt * picV[2] };
        for (int a = 0; a < 3; a++) Xp[a] += inc[a];
        if (fp32_layout)
            for (int a = 0; a < 3; a++) Xp[a] = quantizePos(Xp[a], dx, posBitsOf(h->desc.domain_nodes));
        double incn = inc[0] * inc[0] + inc[1] * inc[1] + inc[2] * inc[2];
        if (incn > dx2) cfl_flag |= 1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double* gv = &gradV[9 * i];
        double* Fp = &h->F[9 * i];
        if (matOf(h, i).model == ZMPM_MODEL_EQUATION_OF_STATE) {
            Fp[0] = (1 + dt * (gv[0] + gv[4] + gv[8])) * Fp[0];
            continue;
        }
        const zmpm_material& mt = matOf(h, i);
        M3 A;
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) A.a[a][c] = (a == c ? 1.0 : 0.0) + dt * gv[3 * a + c];
        M3 Fn = mload(Fp);
        if (mt.fj_mixed == ZMPM_FJ_SOLID) {
            // FJMixedMpmForceHelper::evolveStrain (FJMixedMpmForceHelper.h:168-187): F and
            // J = (1 + dt tr gradV) J, then the J model takes the F model's lambda
            Fn = mmul(A, Fn);
            h->J[i] = (1 + dt * (gv[0] + gv[4] + gv[8])) * h->J[i];
            h->jl[i] = mt.lambda;
        }
        else if (mt.fj_mixed != ZMPM_FJ_FLUID) // a fluid phase keeps F, J and lambda_J
            Fn = mmul(A, Fn);
        // applyPlasticity (MpmSimulationBase.cpp:1517-1529)
        if (mt.plasticity == ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY) vonMisesProject(mt.mu, mt.hardening, h->ys[i], Fn);
        if (mt.plasticity == ZMPM_PLASTICITY_DRUCKER_PRAGER_STVK_HENCKY) druckerPragerProject(mt, h->ys[i], Fn);
        if (mt.plasticity == ZMPM_PLASTICITY_SNOW) snowProject(mt, h->ys[i], Fn);
        mstore(Fn, Fp);
    }
    return cfl_flag;
}

// One explicit substep (MpmSimulationBase::advanceOneTimeStep, MpmSimulationBase.cpp:262-357).
// AnisoFractureSimulation::solveDamageSystem (AnisoFractureSimulation.h:435-612),
// after damageP2G's normalised phase field (:396-433) is on the grid:
//   vol = V0 det F and maxDTilde = max(maxDTilde, D~(F)) per damaged particle
//   (damageP2G :425-432, updateDTilde :284-337);
//   rhs_i = sum_p (dt D~_p / eta + pf_i) vol_p w_ip                      (:457-474)
//   (A x)_i = sum_p vol_p (1 + dt/eta (D~_p + 1)) w_ip x_i
//           + sum_p grad w_ip . (vol_p dt/eta l0^2 sum_j x_j grad w_jp)  (:476-519)
//   Jacobi preconditioner with the first term's diagonal                  (:521-544)
//   ConjugateGradient::solve from x = 0, tolerance 1e-6, 100 iterations  (ConjugateGradient.h:52-103)
//   pf_i <- x_i - pf_i (:603-605), then damageG2P (:614-680):
//   d <- min(max(d, d + sum_i w_ip pf_i), 1), g = (1 - d)^2 (1 - k) + k.
// Only particles whose material carries the phase field take part (the
// reference gives every particle one); a node no such particle reaches has a
// zero diagonal, its preconditioned residual is taken as 0.  Sums in the
// reference's colour order; nodes with m = 0 (idx < 0) are left out.
void solveDamageSystem(zmpm_context* h, double dt)
{
    const int64_t n = h->n;
    const int ng = (int)h->groups.size();
    int64_t N = 0;
    for (const Node& g : h->nodes)
        if (g.idx >= 0) N = std::max<int64_t>(N, g.idx + 1);
    std::vector<char> on((size_t)n, 0);
    std::vector<double> vol((size_t)n, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const zmpm_material& mt = matOf(h, i);
        if (!mt.has_damage || !phaseFieldModel(mt.model)) continue;
        on[i] = 1;
        const M3 F = mload(&h->F[9 * i]);
        vol[i] = h->V0[i] * mdet(F);
        const double dTilde = dTildeOf(mt, F);
        h->dtil[i] = std::max(dTilde, h->dtil[i]);
    }
    // colour sweeps over the damaged particles (deterministic accumulation order)
    auto sweep = [&](auto&& perParticle) {
        for (uint64_t color = 0; color < 8; color++) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int gi = 0; gi < ng; gi++) {
                if ((h->block_offset[gi] & 7) != color) continue;
                for (int idx = h->groups[gi].first; idx <= h->groups[gi].second; idx++) {
                    const int i = h->order[idx];
                    if (on[i]) perParticle(i);
                }
            }
        }
    };
    auto coef = [&](int i) {
        const zmpm_material& mt = matOf(h, i);
        return vol[i] * (1 + ((dt / mt.eta) * (h->dtil[i] + 1)));
    };
    std::vector<double> rhs((size_t)N, 0.0);
    sweep([&](int i) {
        const zmpm_material& mt = matOf(h, i);
        Stencil st;
        makeStencil(h, &h->x[3 * i], st);
        iterateKernel(h, st, [&](const int*, double w, const double*, double, Node& g) {
            if (g.idx < 0) return;
            rhs[(size_t)g.idx] += (((dt * h->dtil[i]) / mt.eta) + g.pf) * vol[i] * w;
        });
    });
    std::vector<double> gradc((size_t)3 * n);
    auto multiply = [&](const std::vector<double>& xv, std::vector<double>& b) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            if (!on[i]) continue;
            const zmpm_material& mt = matOf(h, i);
            double gc[3] = { 0, 0, 0 };
            Stencil st;
            makeStencil(h, &h->x[3 * i], st);
            iterateKernel(h, st, [&](const int*, double, const double* dw, double, Node& g) {
                if (g.idx < 0) return;
                for (int a = 0; a < 3; a++) gc[a] += xv[(size_t)g.idx] * dw[a];
            });
            const double sc = vol[i] * (dt / mt.eta) * mt.l0 * mt.l0;
            for (int a = 0; a < 3; a++) gradc[3 * i + a] = gc[a] * sc;
        }
        std::fill(b.begin(), b.end(), 0.0);
        sweep([&](int i) {
            const double c = coef(i);
            Stencil st;
            makeStencil(h, &h->x[3 * i], st);
            iterateKernel(h, st, [&](const int*, double w, const double* dw, double, Node& g) {
                if (g.idx < 0) return;
                b[(size_t)g.idx] += c * w * xv[(size_t)g.idx];
                b[(size_t)g.idx] += gradc[3 * i] * dw[0] + gradc[3 * i + 1] * dw[1] + gradc[3 * i + 2] * dw[2];
            });
        });
    };
    std::vector<double> diag((size_t)N, 0.0);
    sweep([&](int i) {
        const double c = coef(i);
        Stencil st;
        makeStencil(h, &h->x[3 * i], st);
        iterateKernel(h, st, [&](const int*, double w, const double*, double, Node& g) {
            if (g.idx < 0) return;
            diag[(size_t)g.idx] += c * w;
        });
    });
    auto precondition = [&](const std::vector<double>& in, std::vector<double>& out) {
        for (int64_t q = 0; q < N; q++) out[(size_t)q] = diag[(size_t)q] != 0 ? in[(size_t)q] / diag[(size_t)q] : 0.0;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double r = 0;
        for (int64_t q = 0; q < N; q++) r += a[(size_t)q] * b[(size_t)q];
        return r;
    };
    // ConjugateGradient::solve (ConjugateGradient.h:52-103), relative tolerance
    // 1e-12 of the first residual (double), absolute 1e-6
    std::vector<double> xs((size_t)N, 0.0), r = rhs, q((size_t)N), mr((size_t)N), sdir((size_t)N, 0.0);
    multiply(xs, q);
    for (int64_t k = 0; k < N; k++) r[(size_t)k] -= q[(size_t)k];
    double conv = std::sqrt(dot(r, r));
    const double rel = conv * (h->desc.precision == ZMPM_PREC_FP32 ? 1e-6 : 1e-12);
    double rho_old = std::numeric_limits<double>::max();
    int it = 0;
    for (it = 0;; it++) {
        precondition(r, mr);
        const double rho = dot(mr, r);
        for (int64_t k = 0; k < N; k++) sdir[(size_t)k] = it == 0 ? mr[(size_t)k] : rho / rho_old * sdir[(size_t)k] + mr[(size_t)k];
        multiply(sdir, q);
        const double sq = dot(sdir, q);
        const double alpha = sq != 0 ? rho / sq : std::numeric_limits<double>::max();
        for (int64_t k = 0; k < N; k++) {
            xs[(size_t)k] = alpha * sdir[(size_t)k] + xs[(size_t)k];
            r[(size_t)k] = -alpha * q[(size_t)k] + r[(size_t)k];
        }
        rho_old = rho;
        conv = std::sqrt(dot(r, r));
        if (conv < 1e-6 || conv < rel || it == kDamageCgIterations) break;
    }
    h->cg_iterations = it;
    for (Node& g : h->nodes)
        if (g.idx >= 0) g.pf = xs[(size_t)g.idx] - g.pf;
    // damageG2P
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (!on[i]) continue;
        const zmpm_material& mt = matOf(h, i);
        double pf = 0;
        Stencil st;
        makeStencil(h, &h->x[3 * i], st);
        iterateKernel(h, st, [&](const int*, double w, const double*, double, Node& g) {
            if (g.idx >= 0) pf += w * g.pf;
        });
        double& d = h->d[i];
        const double nd = d + pf;
        d = std::min(std::max(d, nd), 1.0);
        const double k = mt.residual_phase;
        h->g[i] = ((1 - d) * (1 - d) * (1 - k)) + k;
    }
}

// One explicit substep (MpmSimulationBase::advanceOneTimeStep), or its stages
// first..last (zmpm_stage): 0 sortParticlesAndPolluteGrid, 1 forcesUpdateParticleState
// + particlesToGrid with the AnisoMPM damage of its override (the explicit update
// or solveDamageSystem, AnisoFractureSimulation.h:129-141), 2
// gridVelocityExplicitUpdate, 3 (nothing: the damage solve is part of 1 here),
// 4 gridToParticles + evolveStrain + applyPlasticity + the damping passes.
zmpm_status substep(zmpm_context* h, double dt, double t, int first = 0, int last = 4)
{
    if (h->n == 0) return ZMPM_OK;
    if (first <= 0 && 0 <= last) {
        zmpm_status st = sortAndPollute(h);
        if (st != ZMPM_OK) return st;
    }
    const int64_t n = h->n;
    const double dx = h->desc.dx;
    const int ng = (int)h->groups.size();
    const bool apic = h->desc.transfer_scheme == ZMPM_APIC_BLEND_RPIC;
    const bool mls = h->desc.mls_mpm != 0;
    const double D_inverse = 4 / (dx * dx);
    if (first <= 1 && 1 <= last) {

    // forcesUpdateParticleState (MpmSimulationBase.cpp:953-979)
    std::vector<double> vtau((size_t)n * 9);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const zmpm_material& mt = matOf(h, i);
        M3 tau;
        if (mt.fj_mixed && h->fj_state == 1) {
            // FJMixedMpmForceHelper::updateState, SOLVE_J (FJMixedMpmForceHelper.h:106-121):
            // vtau += V0 dpsi/dJ J I, QuadraticVolumePenalty dpsi/dJ = lambda (J - 1)
            const double dpsi_dJ = h->jl[i] * (h->J[i] - 1);
            tau = mzero();
            tau.a[0][0] = tau.a[1][1] = tau.a[2][2] = h->V0[i] * dpsi_dJ * h->J[i];
        }
        else if (mt.fj_mixed) {
            // SOLVE_F (:98-105): the F model's kirchhoff; addFJMixedMpmForce zeroes
            // a fluid's mu and lambda (MpmParticleHandleBase.cpp:818-821)
            zmpm_material fm = mt;
            if (mt.fj_mixed == ZMPM_FJ_FLUID) fm.mu = fm.lambda = 0;
            tau = vtauOf(fm, &h->F[9 * i], h->V0[i], h->g[i]);
        }
        else
            // stress linear in (mu, lambda): a snow particle's hardened moduli scale it
            tau = vtauOf(mt, &h->F[9 * i], h->V0[i] * snowScale(mt, h->ys[i]), h->g[i]);
        mstore(tau, &vtau[9 * i]);
    }

    // particlesToGridWithForceHelper<scheme, deg 2, mls> generic path (MpmSimulationBase.cpp:791-854):
    // C = mass C for APIC_blend_RPIC, 0 for FLIP_blend_PIC; MLS force S (x_i - x_p) w D^-1
    for (uint64_t color = 0; color < 8; color++) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int gi = 0; gi < ng; gi++) {
            if ((h->block_offset[gi] & 7) != color) continue;
            for (int idx = h->groups[gi].first; idx <= h->groups[gi].second; idx++) {
                int i = h->order[idx];
                const double* Xp = &h->x[3 * i];
                double mass = h->m[i];
                double momentum[3] = { mass * h->v[3 * i], mass * h->v[3 * i + 1], mass * h->v[3 * i + 2] };
                double Cm[3][3];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) Cm[r][c] = apic ? mass * h->C[9 * i + 3 * r + c] : 0.0;
                const double* S = &vtau[9 * i];
                Stencil s;
                makeStencil(h, Xp, s);
                iterateKernel(h, s, [&](const int* node, double w, const double* dw, double, Node& g) {
                    double xi_xp[3];
                    for (int a = 0; a < 3; a++) xi_xp[a] = (double)node[a] * dx - Xp[a];
                    double vd[3], fd[3];
                    for (int r = 0; r < 3; r++) {
                        // (velocity_density * [xi-xp; 1]) * w, stress_density * [dw; w] with fp = 0
                        vd[r] = (Cm[r][0] * xi_xp[0] + Cm[r][1] * xi_xp[1] + Cm[r][2] * xi_xp[2] + momentum[r]) * w;
                        if (mls) // (stress_density * [xi - xp; 1]) * (w D^-1), fp = 0
                            fd[r] = (S[3 * r + 0] * xi_xp[0] + S[3 * r + 1] * xi_xp[1] + S[3 * r + 2] * xi_xp[2] + (-0.0) * 1.0)
                                * (w * D_inverse);
                        else
                            fd[r] = S[3 * r + 0] * dw[0] + S[3 * r + 1] * dw[1] + S[3 * r + 2] * dw[2] + (-0.0) * w;
                    }
                    g.m += mass * w;
                    for (int r = 0; r < 3; r++) {
                        g.v[r] += vd[r];
                        g.new_v[r] -= fd[r];
                    }
                });
            }
        }
    }
    // getNumNodes (MpmGrid.h:195-208) and v /= m (MpmSimulationBase.cpp:655-666)
    {
        int64_t cnt = 0;
        for (size_t b = 0; b < h->block_pages.size(); b++)
            for (int e = 0; e < h->NB; e++) {
                Node& g = h->nodes[b * h->NB + e];
                if (g.m != 0) g.idx = cnt++;
            }
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < (int64_t)h->nodes.size(); q++) {
            Node& g = h->nodes[q];
            if (g.idx < 0) continue;
            for (int a = 0; a < 3; a++) g.v[a] /= g.m;
        }
    }

    // AnisoMPM damage (AnisoFractureSimulation.h:129-141): the normalised phase
    // field on the grid (computeLaplacians :150-180 / damageP2G :396-420), then
    // the explicit Laplacian update (updateDamage) or the implicit solve
    if (h->any_damage) {
        for (uint64_t color = 0; color < 8; color++) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int gi = 0; gi < ng; gi++) {
                if ((h->block_offset[gi] & 7) != color) continue;
                for (int idx = h->groups[gi].first; idx <= h->groups[gi].second; idx++) {
                    int i = h->order[idx];
                    double dp = matOf(h, i).has_damage ? h->d[i] : 0.0;
                    Stencil s;
                    makeStencil(h, &h->x[3 * i], s);
                    iterateKernel(h, s, [&](const int*, double w, const double*, double, Node& g) {
                        g.pfm += w;
                        g.pf += dp * w;
                    });
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < (int64_t)h->nodes.size(); q++) {
            Node& g = h->nodes[q];
            if (g.idx >= 0) g.pf /= g.pfm;
        }
    }
    if (h->any_damage && h->desc.implicit_damage)
        solveDamageSystem(h, dt);
    else if (h->any_damage) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            Stencil s;
            makeStencil(h, &h->x[3 * i], s);
            double lap = 0;
            iterateKernel(h, s, [&](const int*, double, const double*, double L, Node& g) {
                if (g.idx >= 0) lap += g.pf * L;
            });
            h->lap[i] = lap;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            const zmpm_material& mt = matOf(h, i);
            if (!mt.has_damage || !phaseFieldModel(mt.model)) continue;
            damageUpdate(mt, mload(&h->F[9 * i]), h->lap[i], dt, h->d[i], h->g[i]);
        }
    }

    }
    // gridVelocityExplicitUpdate (MpmSimulationBase.cpp:1179-1228) with
    // multiObjectCollision (CollisionObject.cpp:66-105) at time t
    if (first <= 2 && 2 <= last) {
        double dvg[3] = { h->desc.gravity[0] * dt, h->desc.gravity[1] * dt, h->desc.gravity[2] * dt };
        const std::vector<Collider>& cols = h->colliders;
        const std::vector<ColliderAt> states = collidersAt(cols, t);
        const int NB = h->NB;
        const Layout L = h->L;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < (int64_t)h->block_pages.size(); b++) {
            for (int e = 0; e < NB; e++) {
                Node& g = h->nodes[(size_t)b * NB + e];
                if (g.idx < 0) continue;
                uint64_t off = (h->block_pages[b] << 12) | ((uint64_t)e << L.data_bits);
                int node[3] = { (int)bitPack(off, L.xmask), (int)bitPack(off, L.ymask), (int)bitPack(off, L.zmask) };
                double xi[3] = { node[0] * dx, node[1] * dx, node[2] * dx };
                if (h->vfield) {
                    // explicit_velocity_field(mi, xi, vi) (MpmSimulationBase.cpp:1187-1195)
                    for (int a = 0; a < 3; a++)
                        g.new_v[a] = h->va[a] + (h->vB[3 * a] * xi[0] + h->vB[3 * a + 1] * xi[1] + h->vB[3 * a + 2] * xi[2]);
                    continue;
                }
                double dt_over_mi = dt / g.m;
                double vi[3];
                for (int a = 0; a < 3; a++) vi[a] = g.v[a] + dvg[a] + g.new_v[a] * dt_over_mi;
                // fext(x, t) (:1208-1212), addExternalBodyForce's
                // f = scale (sin(frequency t) + 1) direction (MpmInitializationHelper.h:1436-1444)
                for (const auto& ff : h->fext) {
                    const double amp = ff.scale * (std::sin(ff.freq * t) + 1);
                    for (int a = 0; a < 3; a++) vi[a] += dt * (amp * ff.dir[a]);
                }
                double old_vi[3] = { vi[0], vi[1], vi[2] };
                multiObjectCollision(cols, states, xi, vi);
                if (h->desc.ignore_collision_objects) // ignoreCollisionObject (:1217-1221)
                    for (int a = 0; a < 3; a++) vi[a] = old_vi[a];
                for (int a = 0; a < 3; a++) g.new_v[a] = vi[a];
            }
        }
    }

    if (!(first <= 4 && 4 <= last)) return ZMPM_OK;
    int cfl_flag = gridToParticlesPass(h, dt, h->desc.apic_rpic_ratio, h->desc.flip_pic_ratio, false);

    // damping passes (MpmSimulationBase.cpp:318-348): grid m, v cleared on the
    // substep's active nodes, particlesToGridHelper (no force), v = mv/m, new_v = v,
    // gridToParticles(0) with apic_rpic_ratio = 0 (RPIC) or flip_pic_ratio = 0 (PIC)
    const int damping = apic ? h->desc.rpic_damping_iteration : h->desc.pic_damping_iteration;
    for (int iter = 0; iter < damping; iter++) {
        for (Node& g : h->nodes)
            if (g.idx >= 0) {
                g.m = 0;
                g.v[0] = g.v[1] = g.v[2] = 0;
            }
        // particlesToGridHelper<scheme, deg 2> (MpmSimulationBase.cpp:860-900) at the
        // moved positions, through the substep's base offsets
        for (uint64_t color = 0; color < 8; color++) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int gi = 0; gi < ng; gi++) {
                if ((h->block_offset[gi] & 7) != color) continue;
                for (int idx = h->groups[gi].first; idx <= h->groups[gi].second; idx++) {
                    int i = h->order[idx];
                    const double* Xp = &h->x[3 * i];
                    double mass = h->m[i];
                    double momentum[3] = { mass * h->v[3 * i], mass * h->v[3 * i + 1], mass * h->v[3 * i + 2] };
                    double Cm[3][3];
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) Cm[r][c] = apic ? mass * h->C[9 * i + 3 * r + c] : 0.0;
                    Stencil s;
                    makeStencil(h, Xp, s);
                    iterateKernel(h, s, [&](const int* node, double w, const double*, double, Node& g) {
                        double xi_xp[3];
                        for (int a = 0; a < 3; a++) xi_xp[a] = (double)node[a] * dx - Xp[a];
                        g.m += mass * w;
                        for (int r = 0; r < 3; r++)
                            g.v[r] += (Cm[r][0] * xi_xp[0] + Cm[r][1] * xi_xp[1] + Cm[r][2] * xi_xp[2] + momentum[r]) * w;
                    }, &h->base_node[3 * i]);
                }
            }
        }
        int64_t cnt = 0;
        for (Node& g : h->nodes)
            if (g.m != 0) g.idx = cnt++;
        for (Node& g : h->nodes) {
            if (g.idx < 0) continue;
            for (int a = 0; a < 3; a++) {
                g.v[a] = g.m != 0 ? g.v[a] / g.m : 0.0;
                g.new_v[a] = g.v[a];
            }
        }
        cfl_flag |= gridToParticlesPass(h, 0.0, 0.0, 0.0, true);
    }
    h->stepped = true;
    if (cfl_flag && h->latched == ZMPM_OK) h->latched = ZMPM_E_CFL;
    return ZMPM_OK;
}

} // namespace

// ================================================================ C ABI
extern "C" {

int32_t zmpm_abi_version(void) { return ZMPM_ABI_VERSION; }
const char* zmpm_build_info(void) { return "zmpm FP64 CPU oracle (test infrastructure)"; }

zmpm_status zmpm_create(const zmpm_scene_desc* desc, zmpm_handle* out)
{
    if (!desc || !out) return ZMPM_E_ARG;
    int dn = desc->domain_nodes;
    if (dn < 64 || dn > 4096 || (dn & (dn - 1)) || !(desc->dx > 0)) return ZMPM_E_ARG;
    if (desc->precision != ZMPM_PREC_FP32 && desc->precision != ZMPM_PREC_FP64) return ZMPM_E_ARG;
    if (desc->transfer_scheme != ZMPM_APIC_BLEND_RPIC && desc->transfer_scheme != ZMPM_FLIP_BLEND_PIC) return ZMPM_E_ARG;
    // MpmSimulationBase::initialize / advanceOneTimeStep assertions (MpmSimulationBase.cpp:108-128, 319, 335)
    const bool apic = desc->transfer_scheme == ZMPM_APIC_BLEND_RPIC;
    if (desc->mls_mpm && !apic) return ZMPM_E_ARG;
    if (apic && !(desc->apic_rpic_ratio >= 0 && desc->apic_rpic_ratio <= 1)) return ZMPM_E_ARG;
    if (desc->rpic_damping_iteration < 0 || desc->pic_damping_iteration < 0) return ZMPM_E_ARG;
    if ((desc->rpic_damping_iteration > 0 && !apic) || (desc->pic_damping_iteration > 0 && apic)) return ZMPM_E_ARG;
    if (desc->implicit_damage != 0 && desc->implicit_damage != 1) return ZMPM_E_ARG;
    auto* h = new zmpm_context();
    h->desc = *desc;
    h->L = makeLayout(desc->precision == ZMPM_PREC_FP32 ? 6 : 7);
    h->NB = 1 << h->L.block_bits;
    *out = h;
    return ZMPM_OK;
}

zmpm_status zmpm_destroy(zmpm_handle h)
{
    delete h;
    return ZMPM_OK;
}

const char* zmpm_last_error(zmpm_handle h) { return h ? h->err.c_str() : "null handle"; }

zmpm_status zmpm_add_body_force(zmpm_handle h, int32_t scene, double scale, double frequency, const double* direction)
{
    if (!h) return ZMPM_E_ARG;
    if (!h->subs.empty()) {
        if (scene < 0 || scene >= (int32_t)h->subs.size()) return fail(h, ZMPM_E_ARG, "no such scene");
        const zmpm_status st = zmpm_add_body_force(h->subs[(size_t)scene], 0, scale, frequency, direction);
        return st == ZMPM_OK ? st : fail(h, st, h->subs[(size_t)scene]->err);
    }
    if (scene != 0) return fail(h, ZMPM_E_ARG, "no such scene");
    if (!direction || !std::isfinite(scale) || !std::isfinite(frequency)) return fail(h, ZMPM_E_ARG, "bad body force");
    if (h->fext.size() >= 4) return fail(h, ZMPM_E_ARG, "at most 4 body forces per scene");
    h->fext.push_back({ scale, frequency, { direction[0], direction[1], direction[2] } });
    return ZMPM_OK;
}

zmpm_status zmpm_set_velocity_field(zmpm_handle h, int32_t scene, const double* a, const double* B)
{
    if (!h) return ZMPM_E_ARG;
    if (!h->subs.empty()) {
        if (scene < 0 || scene >= (int32_t)h->subs.size()) return fail(h, ZMPM_E_ARG, "no such scene");
        const zmpm_status st = zmpm_set_velocity_field(h->subs[(size_t)scene], 0, a, B);
        return st == ZMPM_OK ? st : fail(h, st, h->subs[(size_t)scene]->err);
    }
    if (scene != 0) return fail(h, ZMPM_E_ARG, "no such scene");
    if (a && !B) return fail(h, ZMPM_E_ARG, "the velocity field needs B");
    h->vfield = a != nullptr;
    for (int k = 0; k < 3; k++) h->va[k] = a ? a[k] : 0.0;
    for (int k = 0; k < 9; k++) h->vB[k] = a ? B[k] : 0.0;
    return ZMPM_OK;
}

zmpm_status zmpm_set_fj_state(zmpm_handle h, int32_t state)
{
    if (!h) return ZMPM_E_ARG;
    if (state != 0 && state != 1) return fail(h, ZMPM_E_ARG, "FJ state is 0 (SOLVE_F) or 1 (SOLVE_J)");
    h->fj_state = state;
    for (zmpm_context* c : h->subs) c->fj_state = state;
    return ZMPM_OK;
}

zmpm_status zmpm_set_materials(zmpm_handle h, int32_t count, const zmpm_material* mats)
{
    if (h && h->stage_next) return fail(h, ZMPM_E_ARG, "a substep is in progress (zmpm_stage)");
    if (h && !h->subs.empty()) {
        for (zmpm_context* c : h->subs) {
            const zmpm_status st = zmpm_set_materials(c, count, mats);
            if (st != ZMPM_OK) return fail(h, st, c->err);
        }
        h->mats.assign(mats, mats + count);
        return ZMPM_OK;
    }
    if (!h || count <= 0 || count > 256 || !mats) return fail(h, ZMPM_E_ARG, "bad material table");
    for (int32_t i = 0; i < count; i++) {
        const zmpm_material& m = mats[i];
        if (m.model < 0 || m.model > ZMPM_MODEL_NEO_HOOKEAN_BORDEN) return fail(h, ZMPM_E_ARG, "unknown constitutive model");
        if (m.plasticity < ZMPM_PLASTICITY_NONE || m.plasticity > ZMPM_PLASTICITY_SNOW)
            return fail(h, ZMPM_E_ARG, "unknown plasticity");
        if (m.plasticity && m.model == ZMPM_MODEL_EQUATION_OF_STATE)
            return fail(h, ZMPM_E_ARG, "plasticity needs a model with mu");
        if (m.plasticity == ZMPM_PLASTICITY_SNOW && m.model != ZMPM_MODEL_FIXED_COROTATED && m.model != ZMPM_MODEL_NEO_HOOKEAN)
            return fail(h, ZMPM_E_ARG, "snow plasticity needs a stress linear in mu and lambda (fixed corotated, neo-Hookean)");
        if (m.fj_mixed < ZMPM_FJ_NONE || m.fj_mixed > ZMPM_FJ_FLUID) return fail(h, ZMPM_E_ARG, "unknown FJ-mixed phase");
        if (m.fj_mixed
            && ((m.model != ZMPM_MODEL_FIXED_COROTATED && m.model != ZMPM_MODEL_NEO_HOOKEAN) || m.has_damage
                || m.plasticity == ZMPM_PLASTICITY_SNOW))
            return fail(h, ZMPM_E_ARG,
                        "an FJ-mixed material has a fixed-corotated or neo-Hookean F model, no phase field and no snow plasticity");
    }
    for (int32_t m : h->mat)
        if (m >= count) return fail(h, ZMPM_E_ARG, "material table smaller than the uploaded particles' material indices");
    h->mats.assign(mats, mats + count);
    h->any_damage = false;
    for (auto& m : h->mats) h->any_damage = h->any_damage || m.has_damage;
    return ZMPM_OK;
}

zmpm_status zmpm_add_collider(zmpm_handle h, const zmpm_collider* c, int32_t* id_out)
{
    if (h && !h->subs.empty()) return zmpm_add_collider(h->subs[0], c, id_out);
    if (!h || !c) return ZMPM_E_ARG;
    if (h->colliders.size() >= 16) return fail(h, ZMPM_E_ARG, "too many colliders");
    if (c->shape < ZMPM_SHAPE_HALF_SPACE || c->shape > ZMPM_SHAPE_CAPPED_CYLINDER) return fail(h, ZMPM_E_ARG, "unknown shape");
    if (c->type < ZMPM_STICKY || c->type > ZMPM_GHOST) return fail(h, ZMPM_E_ARG, "unknown collider type");
    if (!((c->scale0 == 0 ? 1.0 : c->scale0) > 0)) return fail(h, ZMPM_E_ARG, "collider scaling must be positive");
    Collider co;
    co.c = *c;
    double nn = std::sqrt(c->normal[0] * c->normal[0] + c->normal[1] * c->normal[1] + c->normal[2] * c->normal[2]);
    for (int a = 0; a < 3; a++) co.n[a] = nn > 0 ? c->normal[a] / nn : 0.0; // HalfSpace ctor normalises
    h->colliders.push_back(co);
    if (id_out) *id_out = (int32_t)h->colliders.size() - 1;
    return ZMPM_OK;
}

zmpm_status zmpm_clear_colliders(zmpm_handle h)
{
    if (h && !h->subs.empty()) {
        for (zmpm_context* c : h->subs) zmpm_clear_colliders(c);
        return ZMPM_OK;
    }
    if (!h) return ZMPM_E_ARG;
    h->colliders.clear();
    return ZMPM_OK;
}

zmpm_status zmpm_upload_batch(zmpm_handle h, const zmpm_particles* p, const int32_t* scene);

zmpm_status zmpm_upload(zmpm_handle h, const zmpm_particles* p)
{
    if (!h || !p || p->n < 0) return ZMPM_E_ARG;
    if (h->stage_next) return fail(h, ZMPM_E_ARG, "a substep is in progress (zmpm_stage)");
    if (!h->subs.empty()) {
        std::vector<int32_t> zero((size_t)p->n, 0);
        return zmpm_upload_batch(h, p, zero.data());
    }
    if (p->n > 0 && (!p->x || !p->v || !p->C || !p->F || !p->m || !p->V0 || !p->material))
        return fail(h, ZMPM_E_ARG, "missing particle array");
    if (h->mats.empty()) return fail(h, ZMPM_E_ARG, "set materials before upload");
    const int64_t n = p->n;
    for (int64_t i = 0; i < n; i++)
        if (p->material[i] < 0 || p->material[i] >= (int32_t)h->mats.size()) return fail(h, ZMPM_E_ARG, "material index out of range");
    h->n = n;
    h->x.assign(p->x, p->x + 3 * n);
    if (h->desc.precision == ZMPM_PREC_FP32)
        for (double& xv : h->x) xv = quantizePos(xv, h->desc.dx, posBitsOf(h->desc.domain_nodes));
    h->v.assign(p->v, p->v + 3 * n);
    h->C.assign(p->C, p->C + 9 * n);
    h->F.assign(p->F, p->F + 9 * n);
    h->m.assign(p->m, p->m + n);
    h->V0.assign(p->V0, p->V0 + n);
    h->mat.assign(p->material, p->material + n);
    h->d.assign(n, 0.0);
    h->g.assign(n, 1.0);
    h->ys.resize(n);
    for (int64_t i = 0; i < n; i++) {
        const zmpm_material& mt = h->mats[(size_t)h->mat[i]];
        h->ys[i] = p->plastic_state ? p->plastic_state[i]
                                    : (mt.plasticity == ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY ? mt.yield_stress
                                       : mt.plasticity == ZMPM_PLASTICITY_SNOW                 ? 1.0
                                                                                               : 0.0);
    }
    h->lap.assign(n, 0.0);
    h->dtil.assign(n, 0.0);
    if (p->max_dtilde) h->dtil.assign(p->max_dtilde, p->max_dtilde + n);
    // addFJMixedMpmForce: J = 1 and the J model's own lambda
    h->J.assign(n, 1.0);
    if (p->J) h->J.assign(p->J, p->J + n);
    h->jl.resize(n);
    for (int64_t i = 0; i < n; i++) h->jl[i] = h->mats[(size_t)h->mat[i]].j_lambda;
    if (p->d) h->d.assign(p->d, p->d + n);
    if (p->g) h->g.assign(p->g, p->g + n);
    h->latched = ZMPM_OK;
    h->stepped = false;
    return ZMPM_OK;
}

zmpm_status batchDownload(zmpm_handle h, zmpm_particles* p)
{
    if (!p || p->n != h->n) return fail(h, ZMPM_E_ARG, "particle count mismatch");
    zmpm_status first = ZMPM_OK;
    std::vector<int64_t> next(h->subs.size(), 0);
    for (size_t s = 0; s < h->subs.size(); s++) {
        zmpm_context* c = h->subs[s];
        if (first == ZMPM_OK && c->latched != ZMPM_OK) first = (zmpm_status)c->latched;
    }
    for (int64_t i = 0; i < h->n; i++) {
        const size_t s = (size_t)h->pscene[(size_t)i];
        const zmpm_context* c = h->subs[s];
        const int64_t k = next[s]++;
        for (int a = 0; a < 3; a++) {
            if (p->x) p->x[3 * i + a] = c->x[3 * k + a];
            if (p->v) p->v[3 * i + a] = c->v[3 * k + a];
        }
        for (int a = 0; a < 9; a++) {
            if (p->C) p->C[9 * i + a] = c->C[9 * k + a];
            if (p->F) p->F[9 * i + a] = c->F[9 * k + a];
        }
        if (p->m) p->m[i] = c->m[k];
        if (p->V0) p->V0[i] = c->V0[k];
        if (p->material) p->material[i] = c->mat[k];
        if (p->d) p->d[i] = c->d[k];
        if (p->g) p->g[i] = c->g[k];
        if (p->laplacian) p->laplacian[i] = c->lap[k];
        if (p->plastic_state) p->plastic_state[i] = c->ys[k];
        if (p->max_dtilde) p->max_dtilde[i] = c->dtil[k];
        if (p->J) p->J[i] = c->J[k];
    }
    return first;
}

// ---- batches: scenes as separate contexts (a batch's scenes never interact)
zmpm_status zmpm_create_batch(const zmpm_scene_desc* descs, int32_t n_scenes, zmpm_handle* out)
{
    if (!descs || !out || n_scenes < 1 || n_scenes > 4096) return ZMPM_E_ARG;
    const zmpm_scene_desc& d0 = descs[0];
    int region = 0;
    for (int32_t s = 0; s < n_scenes; s++) {
        const zmpm_scene_desc& d = descs[s];
        if (d.precision != d0.precision || d.device != d0.device || d.transfer_scheme != d0.transfer_scheme
            || d.flip_pic_ratio != d0.flip_pic_ratio || d.mls_mpm != d0.mls_mpm
            || d.rpic_damping_iteration != d0.rpic_damping_iteration || d.pic_damping_iteration != d0.pic_damping_iteration
            || d.implicit_damage != d0.implicit_damage || !(d.cfl > 0))
            return ZMPM_E_ARG;
        region = std::max(region, d.domain_nodes);
    }
    if (n_scenes > 1 && d0.implicit_damage) return ZMPM_E_ARG;
    int rd = 1;
    while (rd * rd * rd < n_scenes) rd++;
    int grid = 1;
    while (grid < rd * region) grid <<= 1;
    if (grid > 4096) return ZMPM_E_ARG;
    auto* h = new zmpm_context();
    h->desc = d0;
    h->L = makeLayout(d0.precision == ZMPM_PREC_FP32 ? 6 : 7);
    h->NB = 1 << h->L.block_bits;
    for (int32_t s = 0; s < n_scenes; s++) {
        zmpm_handle c = nullptr;
        if (zmpm_create(&descs[s], &c) != ZMPM_OK) {
            delete h;
            return ZMPM_E_ARG;
        }
        h->subs.push_back(c);
    }
    // the regions' layout of the GPU library (page order of each region's first node)
    std::vector<std::pair<uint64_t, int>> order;
    for (int r = 0; r < rd * rd * rd; r++)
        order.push_back({ linearOffset(h->L, (r / (rd * rd)) * region, ((r / rd) % rd) * region, (r % rd) * region), r });
    std::sort(order.begin(), order.end());
    for (int32_t s = 0; s < n_scenes; s++) {
        const int r = order[(size_t)s].second;
        h->origins.push_back({ (r / (rd * rd)) * region, ((r / rd) % rd) * region, (r % rd) * region });
    }
    *out = h;
    return ZMPM_OK;
}

zmpm_status zmpm_batch_origin(zmpm_handle h, int32_t scene, int32_t* origin)
{
    if (!h || !origin) return ZMPM_E_ARG;
    if (h->subs.empty()) {
        if (scene != 0) return ZMPM_E_ARG;
        origin[0] = origin[1] = origin[2] = 0;
        return ZMPM_OK;
    }
    if (scene < 0 || scene >= (int32_t)h->subs.size()) return ZMPM_E_ARG;
    for (int a = 0; a < 3; a++) origin[a] = h->origins[(size_t)scene][a];
    return ZMPM_OK;
}

zmpm_status zmpm_add_scene_collider(zmpm_handle h, int32_t scene, const zmpm_collider* c, int32_t* id_out)
{
    if (!h) return ZMPM_E_ARG;
    if (h->subs.empty()) return scene == 0 ? zmpm_add_collider(h, c, id_out) : fail(h, ZMPM_E_ARG, "no such scene");
    if (scene < 0 || scene >= (int32_t)h->subs.size()) return fail(h, ZMPM_E_ARG, "no such scene");
    return zmpm_add_collider(h->subs[(size_t)scene], c, id_out);
}

zmpm_status zmpm_upload_batch(zmpm_handle h, const zmpm_particles* p, const int32_t* scene)
{
    if (!h || !p || p->n < 0) return ZMPM_E_ARG;
    if (p->n > 0 && !scene) return fail(h, ZMPM_E_ARG, "missing scene array");
    const size_t ns = h->subs.empty() ? 1 : h->subs.size();
    for (int64_t i = 0; i < p->n; i++)
        if (scene[i] < 0 || scene[i] >= (int32_t)ns) return fail(h, ZMPM_E_ARG, "scene index out of range");
    if (h->subs.empty()) return zmpm_upload(h, p);
    h->pscene.assign(scene, scene + p->n);
    h->n = p->n;
    for (size_t s = 0; s < ns; s++) {
        // the scene's particles in upload order
        std::vector<int64_t> idx;
        for (int64_t i = 0; i < p->n; i++)
            if (scene[i] == (int32_t)s) idx.push_back(i);
        const int64_t k = (int64_t)idx.size();
        std::vector<double> x(3 * k), v(3 * k), C(9 * k), F(9 * k), m(k), V0(k), d(k), g(k), ys(k), dt(k), Jv(k, 1.0);
        std::vector<int32_t> mat(k);
        for (int64_t j = 0; j < k; j++) {
            const int64_t i = idx[(size_t)j];
            for (int a = 0; a < 3; a++) {
                x[3 * j + a] = p->x[3 * i + a];
                v[3 * j + a] = p->v[3 * i + a];
            }
            for (int a = 0; a < 9; a++) {
                C[9 * j + a] = p->C[9 * i + a];
                F[9 * j + a] = p->F[9 * i + a];
            }
            m[j] = p->m[i];
            V0[j] = p->V0[i];
            mat[j] = p->material[i];
            if (p->d) d[j] = p->d[i];
            if (p->g) g[j] = p->g[i];
            if (p->plastic_state) ys[j] = p->plastic_state[i];
            if (p->max_dtilde) dt[j] = p->max_dtilde[i];
            if (p->J) Jv[j] = p->J[i];
        }
        zmpm_particles q {};
        q.n = k;
        q.x = x.data();
        q.v = v.data();
        q.C = C.data();
        q.F = F.data();
        q.m = m.data();
        q.V0 = V0.data();
        q.material = mat.data();
        q.d = p->d ? d.data() : nullptr;
        q.g = p->g ? g.data() : nullptr;
        q.plastic_state = p->plastic_state ? ys.data() : nullptr;
        q.max_dtilde = p->max_dtilde ? dt.data() : nullptr;
        q.J = p->J ? Jv.data() : nullptr;
        const zmpm_status st = zmpm_upload(h->subs[s], &q);
        if (st != ZMPM_OK) return fail(h, st, h->subs[s]->err);
    }
    return ZMPM_OK;
}

zmpm_status zmpm_run_batch(zmpm_handle h, int32_t steps, const double* dt, const double* t0)
{
    if (!h || steps < 0 || !dt || !t0) return ZMPM_E_ARG;
    if (h->subs.empty()) return zmpm_run(h, steps, dt[0], t0[0]);
    zmpm_status first = ZMPM_OK;
    for (size_t s = 0; s < h->subs.size(); s++) {
        const zmpm_status st = zmpm_run(h->subs[s], steps, dt[s], t0[s]);
        if (first == ZMPM_OK) first = st;
    }
    return first;
}

zmpm_status zmpm_download(zmpm_handle h, zmpm_particles* p)
{
    if (h && !h->subs.empty()) return batchDownload(h, p);
    if (!h || !p || p->n != h->n) return fail(h, ZMPM_E_ARG, "particle count mismatch");
    const int64_t n = h->n;
    if (p->x) std::copy(h->x.begin(), h->x.end(), p->x);
    if (p->v) std::copy(h->v.begin(), h->v.end(), p->v);
    if (p->C) std::copy(h->C.begin(), h->C.end(), p->C);
    if (p->F) std::copy(h->F.begin(), h->F.end(), p->F);
    if (p->m) std::copy(h->m.begin(), h->m.end(), p->m);
    if (p->V0) std::copy(h->V0.begin(), h->V0.end(), p->V0);
    if (p->material) std::copy(h->mat.begin(), h->mat.end(), p->material);
    if (p->d) std::copy(h->d.begin(), h->d.end(), p->d);
    if (p->g) std::copy(h->g.begin(), h->g.end(), p->g);
    if (p->laplacian) std::copy(h->lap.begin(), h->lap.end(), p->laplacian);
    if (p->plastic_state) std::copy(h->ys.begin(), h->ys.end(), p->plastic_state);
    if (p->max_dtilde) std::copy(h->dtil.begin(), h->dtil.end(), p->max_dtilde);
    if (p->J) std::copy(h->J.begin(), h->J.end(), p->J);
    (void)n;
    zmpm_status s = (zmpm_status)h->latched;
    return s;
}

zmpm_status zmpm_stage(zmpm_handle h, int32_t stage, double dt, double t)
{
    if (!h || stage < ZMPM_STAGE_SORT || stage > ZMPM_STAGE_G2P || !(dt > 0)) return ZMPM_E_ARG;
    if (!h->subs.empty()) return fail(h, ZMPM_E_ARG, "stages run on a single-scene handle");
    if (stage != h->stage_next) return fail(h, ZMPM_E_ARG, "stages run in order: SORT, P2G, GRID_UPDATE, DAMAGE_SOLVE, G2P");
    if (stage == ZMPM_STAGE_SORT) {
        if (h->latched != ZMPM_OK && h->latched != ZMPM_E_CFL) return (zmpm_status)h->latched;
        h->stage_dt = dt;
        h->stage_t = t;
    }
    else if (dt != h->stage_dt || t != h->stage_t)
        return fail(h, ZMPM_E_ARG, "the stages of a substep share its dt and t");
    const zmpm_status st = substep(h, dt, t, stage, stage);
    h->stage_next = stage == ZMPM_STAGE_G2P ? 0 : stage + 1;
    if (st != ZMPM_OK) h->stage_next = 0;
    return st;
}

zmpm_status zmpm_substep(zmpm_handle h, double dt, double t)
{
    if (!h) return ZMPM_E_ARG;
    if (h->stage_next) return fail(h, ZMPM_E_ARG, "a substep is in progress (zmpm_stage)");
    if (!h->subs.empty()) {
        zmpm_status first = ZMPM_OK;
        for (zmpm_context* c : h->subs) {
            const zmpm_status st = zmpm_substep(c, dt, t);
            if (first == ZMPM_OK) first = st;
        }
        return first;
    }
    if (h->latched != ZMPM_OK && h->latched != ZMPM_E_CFL) return (zmpm_status)h->latched;
    return substep(h, dt, t);
}

zmpm_status zmpm_run(zmpm_handle h, int32_t steps, double dt, double t0)
{
    if (!h || steps < 0) return ZMPM_E_ARG;
    if (h->stage_next) return fail(h, ZMPM_E_ARG, "a substep is in progress (zmpm_stage)");
    for (int32_t s = 0; s < steps; s++) {
        zmpm_status st = zmpm_substep(h, dt, t0 + s * dt);
        if (st != ZMPM_OK) return st;
    }
    return ZMPM_OK;
}

zmpm_status zmpm_synchronize(zmpm_handle h)
{
    if (!h) return ZMPM_E_ARG;
    if (!h->subs.empty()) {
        for (zmpm_context* c : h->subs)
            if (c->latched != ZMPM_OK) return (zmpm_status)c->latched;
        return ZMPM_OK;
    }
    return (zmpm_status)h->latched;
}

zmpm_status zmpm_get_binning(zmpm_handle h, int64_t* n_groups, uint64_t* keys, int32_t* order, int32_t* groups, uint64_t* block_offset)
{
    if (h && !h->subs.empty()) return fail(h, ZMPM_E_ARG, "the oracle reports a batch's binning per scene only (a handle per scene)");
    if (!h || !n_groups) return ZMPM_E_ARG;
    *n_groups = (int64_t)h->groups.size();
    if (keys) std::copy(h->keys.begin(), h->keys.end(), keys);
    if (order) std::copy(h->order.begin(), h->order.end(), order);
    if (groups)
        for (size_t gi = 0; gi < h->groups.size(); gi++) {
            groups[2 * gi] = h->groups[gi].first;
            groups[2 * gi + 1] = h->groups[gi].second;
        }
    if (block_offset) std::copy(h->block_offset.begin(), h->block_offset.end(), block_offset);
    return ZMPM_OK;
}

// The grid nodes with mass, ordered by (x, y, z) node coordinates.
struct GridRec {
    int c[3];
    Node* g;
};
std::vector<GridRec> gridRecords(zmpm_context* h)
{
    std::vector<GridRec> recs;
    for (size_t b = 0; b < h->block_pages.size(); b++)
        for (int e = 0; e < h->NB; e++) {
            Node& g = h->nodes[b * h->NB + e];
            if (g.idx < 0) continue;
            uint64_t off = (h->block_pages[b] << 12) | ((uint64_t)e << h->L.data_bits);
            GridRec r { { (int)bitPack(off, h->L.xmask), (int)bitPack(off, h->L.ymask), (int)bitPack(off, h->L.zmask) }, &g };
            recs.push_back(r);
        }
    std::sort(recs.begin(), recs.end(), [](const GridRec& a, const GridRec& b) {
        if (a.c[0] != b.c[0]) return a.c[0] < b.c[0];
        if (a.c[1] != b.c[1]) return a.c[1] < b.c[1];
        return a.c[2] < b.c[2];
    });
    return recs;
}

zmpm_status zmpm_set_grid_velocity(zmpm_handle h, int64_t n_nodes, const double* new_v)
{
    if (!h || !new_v) return ZMPM_E_ARG;
    if (!h->subs.empty()) return fail(h, ZMPM_E_ARG, "stages run on a single-scene handle");
    if (h->stage_next != ZMPM_STAGE_DAMAGE_SOLVE && h->stage_next != ZMPM_STAGE_G2P)
        return fail(h, ZMPM_E_ARG, "the grid velocities are replaced between the GRID_UPDATE and G2P stages");
    std::vector<GridRec> recs = gridRecords(h);
    if (n_nodes != (int64_t)recs.size()) return fail(h, ZMPM_E_ARG, "node count differs from zmpm_get_grid's");
    for (size_t q = 0; q < recs.size(); q++)
        for (int a = 0; a < 3; a++) recs[q].g->new_v[a] = new_v[3 * q + a];
    return ZMPM_OK;
}

zmpm_status zmpm_get_grid(zmpm_handle h, int64_t* n_nodes, int32_t* coords, double* m, double* v, double* new_v, double* pf)
{
    if (!h || !n_nodes) return ZMPM_E_ARG;
    std::vector<GridRec> recs;
    if (!h->subs.empty()) {
        // a batch: every scene's nodes at its region's origin in the handle's grid
        for (size_t sc = 0; sc < h->subs.size(); sc++) {
            std::vector<GridRec> r = gridRecords(h->subs[sc]);
            for (GridRec& g : r)
                for (int a = 0; a < 3; a++) g.c[a] += h->origins[sc][a];
            recs.insert(recs.end(), r.begin(), r.end());
        }
        std::sort(recs.begin(), recs.end(), [](const GridRec& a, const GridRec& b) {
            if (a.c[0] != b.c[0]) return a.c[0] < b.c[0];
            if (a.c[1] != b.c[1]) return a.c[1] < b.c[1];
            return a.c[2] < b.c[2];
        });
    }
    else
        recs = gridRecords(h);
    const bool any_damage = h->subs.empty() ? h->any_damage : h->subs[0]->any_damage;
    *n_nodes = (int64_t)recs.size();
    for (size_t q = 0; q < recs.size(); q++) {
        const Node& g = *recs[q].g;
        if (coords)
            for (int a = 0; a < 3; a++) coords[3 * q + a] = recs[q].c[a];
        if (m) m[q] = g.m;
        if (v)
            for (int a = 0; a < 3; a++) v[3 * q + a] = g.v[a];
        if (new_v)
            for (int a = 0; a < 3; a++) new_v[3 * q + a] = g.new_v[a];
        if (pf) pf[q] = any_damage ? g.pf : 0.0;
    }
    return ZMPM_OK;
}

zmpm_status zmpm_set_timing(zmpm_handle h, int32_t)
{
    return h ? ZMPM_OK : ZMPM_E_ARG;
}

// calculateDt (MpmSimulationBase.cpp:1127-1151) over evalMaxParticleSpeed (:1613-1638)
zmpm_status zmpm_calculate_dt(zmpm_handle h, double max_dt, double* dt, double* max_speed, double* min_corner,
                              double* max_corner)
{
    if (h && !h->subs.empty()) {
        // the smallest scene step, the largest speed and the box over the scenes' frames
        if (!(max_dt > 0)) return ZMPM_E_ARG;
        double best = max_dt, sp = 0, lo[3] = { 3.4028234663852886e38, 3.4028234663852886e38, 3.4028234663852886e38 },
               hi[3] = { -3.4028234663852886e38, -3.4028234663852886e38, -3.4028234663852886e38 };
        bool any = false;
        for (zmpm_context* c : h->subs) {
            if (c->n == 0) continue;
            double d, s1, l[3], u[3];
            const zmpm_status st = zmpm_calculate_dt(c, max_dt, &d, &s1, l, u);
            if (st != ZMPM_OK) return st;
            if (s1 != 0) {
                best = any ? std::min(best, d) : d;
                any = true;
            }
            sp = std::max(sp, s1);
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], l[a]);
                hi[a] = std::max(hi[a], u[a]);
            }
        }
        if (dt) *dt = best;
        if (max_speed) *max_speed = sp;
        for (int a = 0; a < 3; a++) {
            if (min_corner) min_corner[a] = lo[a];
            if (max_corner) max_corner[a] = hi[a];
        }
        return ZMPM_OK;
    }
    if (!h || !(max_dt > 0)) return ZMPM_E_ARG;
    const double lowest = (double)std::numeric_limits<float>::lowest();
    double r[7] = { 0.0, lowest, lowest, lowest, lowest, lowest, lowest };
    for (int64_t i = 0; i < h->n; i++) {
        const double* v = &h->v[3 * i];
        r[0] = std::max(r[0], std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        for (int a = 0; a < 3; a++) {
            r[1 + a] = std::max(r[1 + a], h->x[3 * i + a]);
            r[4 + a] = std::max(r[4 + a], -h->x[3 * i + a]);
        }
    }
    if (max_speed) *max_speed = r[0];
    for (int a = 0; a < 3; a++) {
        if (max_corner) max_corner[a] = r[1 + a];
        if (min_corner) min_corner[a] = -r[4 + a];
    }
    if (dt) *dt = r[0] != 0 ? h->desc.cfl * h->desc.dx / r[0] : max_dt;
    return ZMPM_OK;
}

// The allocation-free variant has no device here: `out` is host memory of 8
// doubles (dt, max speed, min corner[3], max corner[3]); reduction_wgs is
// checked against the same limit.
zmpm_status zmpm_calculate_dt_noalloc(zmpm_handle h, double max_dt, int32_t reduction_wgs, double* out)
{
    if (!h || !(max_dt > 0) || !out) return ZMPM_E_ARG;
    if (reduction_wgs < 1 || reduction_wgs > 1024) return fail(h, ZMPM_E_ARG, "reduction_wgs must be in [1, 1024]");
    return zmpm_calculate_dt(h, max_dt, &out[0], &out[1], &out[2], &out[5]);
}

zmpm_status zmpm_get_timings(zmpm_handle h, int32_t* count, const char**, double*)
{
    if (!h || !count) return ZMPM_E_ARG;
    *count = 0;
    return ZMPM_OK;
}

// Launch configuration: the oracle has no launches; it keeps and returns the
// values it is given (all ones by default) so callers can run unchanged.
zmpm_status zmpm_get_launch_config(zmpm_handle h, zmpm_launch_config* cfg)
{
    if (!h || !cfg) return ZMPM_E_ARG;
    *cfg = h->launch;
    return ZMPM_OK;
}

zmpm_status zmpm_set_launch_config(zmpm_handle h, const zmpm_launch_config* cfg)
{
    if (!h || !cfg) return ZMPM_E_ARG;
    if (cfg->p2g_wgs < 1 || cfg->grid_wgs < 1 || cfg->prep_wgs < 1 || cfg->p2g_split < 1)
        return ZMPM_E_ARG;
    h->launch = *cfg;
    return ZMPM_OK;
}

zmpm_status zmpm_autotune(zmpm_handle h, double dt, double, int32_t steps, int32_t, zmpm_launch_config* cfg)
{
    if (!h || !(dt > 0) || steps < 1) return ZMPM_E_ARG;
    if (cfg) *cfg = h->launch;
    return ZMPM_OK;
}

// The oracle has nothing to tune or cache: the configuration it keeps, never
// from a cache file.
zmpm_status zmpm_autotune_cached(zmpm_handle h, double dt, double t0, int32_t steps, int32_t tune_variants,
                                 const char* cache_path, zmpm_launch_config* cfg, int32_t* from_cache)
{
    if (!cache_path) return ZMPM_E_ARG;
    if (from_cache) *from_cache = 0;
    return zmpm_autotune(h, dt, t0, steps, tune_variants, cfg);
}

// ------------------------------------------------ unit entry points (pinning)
static Collider makeCollider(const zmpm_collider& c)
{
    Collider co;
    co.c = c;
    const double nn = std::sqrt(c.normal[0] * c.normal[0] + c.normal[1] * c.normal[1] + c.normal[2] * c.normal[2]);
    for (int a = 0; a < 3; a++) co.n[a] = nn > 0 ? c.normal[a] / nn : 0.0; // HalfSpace ctor normalises
    return co;
}

// multiObjectCollision of one grid node at x, time t; v resolved in place.
// Returns 1 when any object collided.
int32_t zmpm_oracle_collide(const zmpm_collider* cols, int32_t count, const double* x, double t, double* v)
{
    std::vector<Collider> cs;
    for (int i = 0; i < count; i++) cs.push_back(makeCollider(cols[i]));
    return multiObjectCollision(cs, collidersAt(cs, t), x, v) ? 1 : 0;
}

// AnalyticCollisionObject::queryInside (CollisionObject.cpp:234-249) of one
// object at world point x, time t: inside flag, phi scaled by s, world normal R N.
int32_t zmpm_oracle_query(const zmpm_collider* c, const double* x, double t, double* phi, double* n)
{
    const Collider co = makeCollider(*c);
    const ColliderAt st = colliderAt(*c, t);
    double xb[3], X[3], N[3];
    for (int a = 0; a < 3; a++) xb[a] = x[a] - st.b[a];
    for (int a = 0; a < 3; a++) X[a] = (st.R.a[0][a] * xb[0] + st.R.a[1][a] * xb[1] + st.R.a[2][a] * xb[2]) / st.s;
    const bool inside = levelSetInside(co, X, *phi, N);
    *phi *= st.s;
    for (int a = 0; a < 3; a++) n[a] = st.R.a[a][0] * N[0] + st.R.a[a][1] * N[1] + st.R.a[a][2] * N[2];
    return inside ? 1 : 0;
}

// OpenMP threads of the oracle's parallel loops (results do not depend on it)
void zmpm_oracle_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t zmpm_oracle_get_threads(void) { return omp_get_max_threads(); }

// CG iterations of the handle's last implicit damage solve (-1: none yet)
int32_t zmpm_oracle_cg_iterations(zmpm_handle h) { return h ? h->cg_iterations : -1; }

uint64_t zmpm_oracle_linear_offset(int32_t log2_struct, int32_t i, int32_t j, int32_t k)
{
    return linearOffset(makeLayout(log2_struct), i, j, k);
}

void zmpm_oracle_masks(int32_t log2_struct, uint64_t* xyz)
{
    Layout L = makeLayout(log2_struct);
    xyz[0] = L.xmask;
    xyz[1] = L.ymask;
    xyz[2] = L.zmask;
}

void zmpm_oracle_bspline(double x, int32_t* base, double* w, double* dw, double* ddw)
{
    int b;
    bspline(x, b, w, dw, ddw);
    *base = b;
}

// DruckerPragerStvkHencky::projectStrain on one deformation gradient (row-major, in place)
void zmpm_oracle_drucker_prager(const zmpm_material* mat, double* logJp, double* F)
{
    M3 f = mload(F);
    druckerPragerProject(*mat, *logJp, f);
    mstore(f, F);
}

// VonMisesStvkHencky::projectStrain on one deformation gradient (row-major, in place)
void zmpm_oracle_snow(const zmpm_material* mat, double* Jp, double* F)
{
    M3 f = mload(F);
    snowProject(*mat, *Jp, f);
    mstore(f, F);
}

void zmpm_oracle_von_mises(double mu, double xi, double* yield_stress, double* F)
{
    M3 f = mload(F);
    vonMisesProject(mu, xi, *yield_stress, f);
    mstore(f, F);
}

int32_t zmpm_oracle_svd(const double* A, double* U, double* sigma, double* V)
{
    M3 a = mload(A), u, v;
    int c = svd3(a, u, sigma, v, kSvdTol);
    mstore(u, U);
    mstore(v, V);
    return c;
}

void zmpm_oracle_first_piola(const zmpm_material* mat, const double* F, double g, double* P)
{
    mstore(firstPiola(*mat, mload(F), g), P);
}

double zmpm_oracle_psi(const zmpm_material* mat, const double* F, double g)
{
    return psiOf(*mat, mload(F), g);
}

void zmpm_oracle_vtau(const zmpm_material* mat, const double* F, double V0, double g, double* vtau)
{
    mstore(vtauOf(*mat, F, V0, g), vtau);
}

void zmpm_oracle_damage(const zmpm_material* mat, const double* F, double lap, double dt, double* d, double* g)
{
    damageUpdate(*mat, mload(F), lap, dt, *d, *g);
}

void zmpm_oracle_sym_eigen(const double* S, double* lam, double* vecs)
{
    M3 vv;
    symEigen(mload(S), lam, vv);
    mstore(vv, vecs);
}

} // extern "C"
