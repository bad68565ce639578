// zmpm_math.hpp -- device math of the explicit MPM substep, templated on the
// particle scalar T (float for the FP32 form, double for the FP64 plain form).
//
// Each routine names the reference lines whose behaviour it reproduces.  The
// 3x3 SVD is the implicit-shift symmetric QR method of Gast, Fu, Jiang and
// Teran (UCLA CAM report 16-04, 2016) that Ziran uses
// (Lib/Ziran/Math/Linear/ImplicitQRSVD.h:354-517); its sign and ordering
// conventions define U, sigma and V, so the steps follow the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zmpm {

template <class T> struct Lim;
template <> struct Lim<float> {
    static constexpr float eps = 1.1920928955078125e-7f;
    static constexpr float svd_tol = 128 * eps; // ImplicitQRSVD.h:518-525
};
template <> struct Lim<double> {
    static constexpr double eps = 2.220446049250313e-16;
    static constexpr double svd_tol = 1024 * eps; // ImplicitQRSVD.h:526-533
};

// Square root, reciprocal square root and division.  The float form uses the
// hardware v_sqrt/v_rsq/v_rcp (about 1 ulp), well inside its error budget;
// the double form keeps correctly rounded operations (parity with the FP64
// oracle).
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float drsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double drsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float ddiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }
__device__ __forceinline__ float dabs(float x) { return fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return fabs(x); }
__device__ __forceinline__ float dcopysign(float a, float b) { return copysignf(a, b); }
__device__ __forceinline__ double dcopysign(double a, double b) { return copysign(a, b); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dpow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double dpow(double a, double b) { return pow(a, b); }
// Products and differences that must not be fused into an FMA (base-node
// flooring, MpmSimulationBase.cpp:1545 and BSplines.h:16-20).
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float sub_rn(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }

// --------------------------------------------------------------- 3x3 matrices
template <class T>
struct Mat3 {
    T a[3][3];
    __device__ __forceinline__ T& operator()(int i, int j) { return a[i][j]; }
    __device__ __forceinline__ const T& operator()(int i, int j) const { return a[i][j]; }
};

template <class T>
__device__ __forceinline__ void setIdentity(Mat3<T>& M)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M.a[i][j] = (i == j) ? T(1) : T(0);
}

template <class T>
__device__ __forceinline__ Mat3<T> mul(const Mat3<T>& A, const Mat3<T>& B)
{
    Mat3<T> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
    return r;
}

// A * B^T
template <class T>
__device__ __forceinline__ Mat3<T> mulT(const Mat3<T>& A, const Mat3<T>& B)
{
    Mat3<T> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.a[i][j] = A.a[i][0] * B.a[j][0] + A.a[i][1] * B.a[j][1] + A.a[i][2] * B.a[j][2];
    return r;
}

template <class T>
__device__ __forceinline__ T det3(const Mat3<T>& m)
{
    return m.a[0][0] * (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1])
        - m.a[1][0] * (m.a[0][1] * m.a[2][2] - m.a[0][2] * m.a[2][1])
        + m.a[2][0] * (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]);
}

// J F^{-T} (cofactorMatrix, DenseExt.h:240-252)
template <class T>
__device__ __forceinline__ Mat3<T> cofactor(const Mat3<T>& Fm)
{
    const T(*F)[3] = Fm.a;
    Mat3<T> A;
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

// ------------------------------------------------------- Givens rotations
// Conventions of Givens.h:32-181: c,s with (c -s; s c)(a; b) = (*; 0).
template <class T>
struct Giv {
    T c, s;
};

template <class T>
__device__ __forceinline__ Giv<T> givens(T a, T b) // Givens.h:72-85
{
    Giv<T> g { T(1), T(0) };
    T d = a * a + b * b;
    if (d != T(0)) {
        T t = drsqrt(d);
        g.c = a * t;
        g.s = -b * t;
    }
    return g;
}

template <class T>
__device__ __forceinline__ Giv<T> givensUnconventional(T a, T b) // Givens.h:92-105
{
    Giv<T> g { T(0), T(1) };
    T d = a * a + b * b;
    if (d != T(0)) {
        T t = drsqrt(d);
        g.s = a * t;
        g.c = b * t;
    }
    return g;
}

template <class T, int I, int K>
__device__ __forceinline__ void rowRot(Mat3<T>& A, Giv<T> g) // Givens.h:131-140
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T t1 = A.a[I][j], t2 = A.a[K][j];
        A.a[I][j] = g.c * t1 - g.s * t2;
        A.a[K][j] = g.s * t1 + g.c * t2;
    }
}

template <class T, int I, int K>
__device__ __forceinline__ void colRot(Mat3<T>& A, Giv<T> g) // Givens.h:150-159
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T t1 = A.a[j][I], t2 = A.a[j][K];
        A.a[j][I] = g.c * t1 - g.s * t2;
        A.a[j][K] = g.s * t1 + g.c * t2;
    }
}

template <class T>
__device__ __forceinline__ void zeroChase(Mat3<T>& H, Mat3<T>& U, Mat3<T>& V) // Givens.h:195-240
{
    Giv<T> r1 = givens(H.a[0][0], H.a[1][0]);
    Giv<T> r2;
    if (H.a[1][0] != T(0))
        r2 = givens(H.a[0][0] * H.a[0][1] + H.a[1][0] * H.a[1][1], H.a[0][0] * H.a[0][2] + H.a[1][0] * H.a[1][2]);
    else
        r2 = givens(H.a[0][1], H.a[0][2]);
    rowRot<T, 0, 1>(H, r1);
    colRot<T, 1, 2>(H, r2);
    colRot<T, 1, 2>(V, r2);
    Giv<T> r3 = givens(H.a[1][1], H.a[2][1]);
    rowRot<T, 1, 2>(H, r3);
    colRot<T, 0, 1>(U, r1);
    colRot<T, 1, 2>(U, r3);
}

template <class T>
__device__ __forceinline__ T wilkinsonShift(T a1, T b1, T a2) // ImplicitQRSVD.h:260-272
{
    T d = T(0.5) * (a1 - a2);
    T bs = b1 * b1;
    return a2 - dcopysign(bs / (dabs(d) + dsqrt(d * d + bs)), d);
}

// 2x2 SVD of B[t..t+1][t..t+1] (ImplicitQRSVD.h:45-66, 116-175), then the
// process<t> bookkeeping (:277-293).
template <class T, int t>
__device__ __forceinline__ void process(const Mat3<T>& B, Mat3<T>& U, T* sigma, Mat3<T>& V)
{
    constexpr int other = (t == 1) ? 0 : 2;
    sigma[other] = B.a[other][other];
    T A00 = B.a[t][t], A01 = B.a[t][t + 1], A10 = B.a[t + 1][t], A11 = B.a[t + 1][t + 1];
    T x0 = A00 + A11, x1 = A10 - A01;
    T den = dsqrt(x0 * x0 + x1 * x1);
    Giv<T> u { T(1), T(0) };
    if (den != T(0)) {
        u.c = x0 / den;
        u.s = -x1 / den;
    }
    T S00 = u.c * A00 - u.s * A10;
    T S01 = u.c * A01 - u.s * A11;
    T S11 = u.s * A01 + u.c * A11;
    T cosine, sine, s0, s1;
    T x = S00, y = S01, z = S11;
    T y2 = y * y;
    if (y2 == T(0)) {
        cosine = T(1);
        sine = T(0);
        s0 = x;
        s1 = z;
    }
    else {
        T tau = T(0.5) * (x - z);
        T w = dsqrt(tau * tau + y2);
        T tt = (tau > T(0)) ? y / (tau + w) : y / (tau - w);
        cosine = T(1) / dsqrt(tt * tt + T(1));
        sine = -tt * cosine;
        T c2 = cosine * cosine;
        T csy = T(2) * cosine * sine * y;
        T s2 = sine * sine;
        s0 = c2 * x - csy + s2 * z;
        s1 = s2 * x + csy + c2 * z;
    }
    Giv<T> v;
    if (s0 < s1) {
        T tmp = s0;
        s0 = s1;
        s1 = tmp;
        v.c = -sine;
        v.s = cosine;
    }
    else {
        v.c = cosine;
        v.s = sine;
    }
    Giv<T> uu { u.c * v.c - u.s * v.s, u.s * v.c + u.c * v.s };
    sigma[t] = s0;
    sigma[t + 1] = s1;
    colRot<T, t, t + 1>(U, uu);
    colRot<T, t, t + 1>(V, v);
}

template <class T>
__device__ __forceinline__ void flipSign(int i, Mat3<T>& U, T* sigma)
{
    sigma[i] = -sigma[i];
#pragma unroll
    for (int r = 0; r < 3; r++) U.a[r][i] = -U.a[r][i];
}

template <class T, int I, int J>
__device__ __forceinline__ void swapCols(Mat3<T>& A)
{
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T t = A.a[r][I];
        A.a[r][I] = A.a[r][J];
        A.a[r][J] = t;
    }
}

template <class T, int I>
__device__ __forceinline__ void negCol(Mat3<T>& A)
{
#pragma unroll
    for (int r = 0; r < 3; r++) A.a[r][I] = -A.a[r][I];
}

template <class T>
__device__ __forceinline__ void sort0(Mat3<T>& U, T* sigma, Mat3<T>& V) // ImplicitQRSVD.h:308-341
{
    if (dabs(sigma[1]) >= dabs(sigma[2])) {
        if (sigma[1] < T(0)) {
            flipSign(1, U, sigma);
            flipSign(2, U, sigma);
        }
        return;
    }
    if (sigma[2] < T(0)) {
        flipSign(1, U, sigma);
        flipSign(2, U, sigma);
    }
    T t = sigma[1];
    sigma[1] = sigma[2];
    sigma[2] = t;
    swapCols<T, 1, 2>(U);
    swapCols<T, 1, 2>(V);
    if (sigma[1] > sigma[0]) {
        t = sigma[0];
        sigma[0] = sigma[1];
        sigma[1] = t;
        swapCols<T, 0, 1>(U);
        swapCols<T, 0, 1>(V);
    }
    else {
        negCol<T, 2>(U);
        negCol<T, 2>(V);
    }
}

template <class T>
__device__ __forceinline__ void sort1(Mat3<T>& U, T* sigma, Mat3<T>& V) // second sort<1> overload
{
    if (dabs(sigma[0]) >= sigma[1]) {
        if (sigma[0] < T(0)) {
            flipSign(0, U, sigma);
            flipSign(2, U, sigma);
        }
        return;
    }
    T t = sigma[0];
    sigma[0] = sigma[1];
    sigma[1] = t;
    swapCols<T, 0, 1>(U);
    swapCols<T, 0, 1>(V);
    if (dabs(sigma[1]) < dabs(sigma[2])) {
        t = sigma[1];
        sigma[1] = sigma[2];
        sigma[2] = t;
        swapCols<T, 1, 2>(U);
        swapCols<T, 1, 2>(V);
    }
    else {
        negCol<T, 1>(U);
        negCol<T, 1>(V);
    }
    if (sigma[1] < T(0)) {
        flipSign(1, U, sigma);
        flipSign(2, U, sigma);
    }
}

// A = U diag(sigma) V^T with U, V rotations; sigma sorted by decreasing
// magnitude, only sigma[2] may be negative.  Returns the QR sweep count
// (bounded: the loop also stops after 64 sweeps so every lane terminates).
template <class T>
__device__ __forceinline__ int svd3(const Mat3<T>& A, Mat3<T>& U, T* sigma, Mat3<T>& V)
{
    Mat3<T> B = A;
    setIdentity(U);
    setIdentity(V);
    { // makeUpperBidiag (Givens.h:252-271)
        Giv<T> r = givens(B.a[1][0], B.a[2][0]);
        rowRot<T, 1, 2>(B, r);
        colRot<T, 1, 2>(U, r);
        zeroChase(B, U, V);
    }
    int count = 0;
    T alpha_1 = B.a[0][0], beta_1 = B.a[0][1], alpha_2 = B.a[1][1], alpha_3 = B.a[2][2], beta_2 = B.a[1][2];
    T gamma_1 = alpha_1 * beta_1, gamma_2 = alpha_2 * beta_2;
    T tol = Lim<T>::svd_tol;
    T sc = T(0.5) * dsqrt(alpha_1 * alpha_1 + alpha_2 * alpha_2 + alpha_3 * alpha_3 + beta_1 * beta_1 + beta_2 * beta_2);
    tol *= (sc > T(1) ? sc : T(1));
    while (dabs(beta_2) > tol && dabs(beta_1) > tol && dabs(alpha_1) > tol && dabs(alpha_2) > tol
        && dabs(alpha_3) > tol && count < 64) {
        T mu = wilkinsonShift(alpha_2 * alpha_2 + beta_1 * beta_1, gamma_2, alpha_3 * alpha_3 + beta_2 * beta_2);
        Giv<T> r = givens(alpha_1 * alpha_1 - mu, gamma_1);
        colRot<T, 0, 1>(B, r);
        colRot<T, 0, 1>(V, r);
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
    if (dabs(beta_2) <= tol) {
        process<T, 0>(B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else if (dabs(beta_1) <= tol) {
        process<T, 1>(B, U, sigma, V);
        sort1(U, sigma, V);
    }
    else if (dabs(alpha_2) <= tol) {
        Giv<T> r1 = givensUnconventional(B.a[1][2], B.a[2][2]);
        rowRot<T, 1, 2>(B, r1);
        colRot<T, 1, 2>(U, r1);
        process<T, 0>(B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else if (dabs(alpha_3) <= tol) {
        Giv<T> r1 = givens(B.a[1][1], B.a[1][2]);
        colRot<T, 1, 2>(B, r1);
        colRot<T, 1, 2>(V, r1);
        Giv<T> r2 = givens(B.a[0][0], B.a[0][2]);
        colRot<T, 0, 2>(B, r2);
        colRot<T, 0, 2>(V, r2);
        process<T, 0>(B, U, sigma, V);
        sort0(U, sigma, V);
    }
    else {
        // alpha_1 ~ 0 (also reached if the sweep cap hits, which leaves B nearly diagonal)
        Giv<T> r1 = givensUnconventional(B.a[0][1], B.a[1][1]);
        rowRot<T, 0, 1>(B, r1);
        colRot<T, 0, 1>(U, r1);
        Giv<T> r2 = givensUnconventional(B.a[0][2], B.a[2][2]);
        rowRot<T, 0, 2>(B, r2);
        colRot<T, 0, 2>(U, r2);
        process<T, 1>(B, U, sigma, V);
        sort1(U, sigma, V);
    }
    return count;
}

// Rotation of the polar decomposition, R = U V^T (ImplicitQRSVD.h:546-558).
// The float form switches to double when the two smallest signed singular
// values nearly cancel (F close to inversion), where R is ill-conditioned.
template <class T>
__device__ __forceinline__ Mat3<T> polarRotation(const Mat3<T>& F)
{
    Mat3<T> U, V;
    T s[3];
    svd3(F, U, s, V);
    return mulT(U, V);
}

// Float form: the same rotation from the scaled Newton iteration
// X <- (g X + X^{-T} / g) / 2, g = |det X|^{-1/3} (Higham, "Computing the polar
// decomposition -- with applications", SIAM J. Sci. Stat. Comput. 7, 1986),
// which converges quadratically to U V^T for det F > 0 without the branchy QR
// sweeps.  Near inversion (small or negative det relative to |F|^3, where R is
// ill-conditioned or the orthogonal factor is a reflection) the implicit-QR
// SVD runs in double.
__device__ __forceinline__ Mat3<float> polarRotationSvd64(const Mat3<float>& F)
{
    Mat3<double> Fd, Ud, Vd;
    double sd[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Fd.a[i][j] = F.a[i][j];
    svd3(Fd, Ud, sd, Vd);
    Mat3<double> Rd = mulT(Ud, Vd);
    Mat3<float> R;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) R.a[i][j] = (float)Rd.a[i][j];
    return R;
}

template <>
__device__ __forceinline__ Mat3<float> polarRotation<float>(const Mat3<float>& F)
{
    float n2 = 0.f;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) n2 += F.a[i][j] * F.a[i][j];
    const float J = det3(F);
    // det F >= 1e-3 (|F|_F^2 / 3)^{3/2}: smallest singular value bounded away from 0
    const float ref = n2 * (1.f / 3.f);
    bool ok = J > 1e-3f * ref * sqrtf(ref);
    Mat3<float> X = F;
    if (ok) {
        int it = 0;
        for (; it < 12; it++) {
            Mat3<float> C = cofactor(X);
            const float d = X.a[0][0] * C.a[0][0] + X.a[0][1] * C.a[0][1] + X.a[0][2] * C.a[0][2];
            const float g = __builtin_amdgcn_exp2f(-__builtin_amdgcn_logf(fabsf(d)) * (1.f / 3.f));
            const float a = 0.5f * g, b = 0.5f * __builtin_amdgcn_rcpf(g * d);
            float delta = 0.f;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float xn = a * X.a[i][j] + b * C.a[i][j];
                    const float e = xn - X.a[i][j];
                    delta += e * e;
                    X.a[i][j] = xn;
                }
            if (delta < 1e-12f) break;
        }
        ok = it < 12;
    }
    return ok ? X : polarRotationSvd64(F);
}

// ------------------------------------------------------ constitutive models
// Fixed corotated Kirchhoff stress tau = P F^T with P = 2 mu (F - R) + lambda (J-1) J F^{-T}
// (CorotatedElasticity.cpp:55-76, HyperelasticConstitutiveModel.h:75-80).
template <class T>
__device__ __forceinline__ Mat3<T> corotatedTau(const Mat3<T>& F, T mu, T lambda)
{
    T J = det3(F);
    Mat3<T> JFinvT = cofactor(F);
    Mat3<T> R = polarRotation(F);
    Mat3<T> P;
    T two_mu = T(2) * mu, lj = lambda * (J - T(1));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) P.a[i][j] = two_mu * (F.a[i][j] - R.a[i][j]) + lj * JFinvT.a[i][j];
    return mulT(P, F);
}

// QRStableNeoHookean Kirchhoff stress tau = P F^T, symmetric (xx xy xz yy yz zz).
// The reference evaluates P = Q A R^{-T} from the Givens QR F = Q R, with A the
// symmetric matrix whose upper triangle is dPsiHat/dR R^T
// (QRStableNeoHookean.h:108-213, QrBasedHelper::evaluateP, QrBasedHelper.h:28-40).
// Then tau = Q A R^{-T} R^T Q^T = Q A Q^T, and with the entries of dPsiHat/dR
// (upper triangular: g mu R_ij off the diagonal; (vol - g mu J) / R_ii + g mu R_ii
// on it, vol = g_l lambda (J - 1) J, g_l = g for J >= 1 else 1; fibre terms
// kx (R_00 - 1) on (0,0) and ky (s - 1) R_i1 / s on (0,1), (1,1), s = |(R_01, R_11)|)
//   A = (vol - g mu J) I + g mu R R^T + kx (R_00 - 1) R_00 e0 e0^T + c r r^T,
//   r = (R_01, R_11, 0), c = ky (s - 1) / s.
// Since Q R R^T Q^T = F F^T, Q e0 R_00 = f0 and Q r = f1 (the first two columns
// of F, |f0| = R_00 >= 0 and |f1| = s under the sign convention, det Q = 1 so
// J = R_00 R_11 R_22 = det F):
//   tau = g mu F F^T + (vol - g mu J) I + kx (|f0| - 1) / |f0| f0 f0^T
//         + ky (|f1| - 1) / |f1| f1 f1^T.
// Same stress as the reference up to rounding, without the QR sweep.
template <class T>
__device__ __forceinline__ void qrStableTau(const Mat3<T>& F, T mu, T lambda, T g, T kx, T ky, T* tau, T* Jout)
{
    const T J = det3(F);
    const T gl = (J >= T(1)) ? g * lambda : lambda;
    const T gm = g * mu;
    const T iso = gl * (J - T(1)) * J - gm * J;
    const T r0[3] = { F.a[0][0], F.a[0][1], F.a[0][2] };
    const T r1[3] = { F.a[1][0], F.a[1][1], F.a[1][2] };
    const T r2[3] = { F.a[2][0], F.a[2][1], F.a[2][2] };
    tau[0] = gm * (r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]) + iso;
    tau[1] = gm * (r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2]);
    tau[2] = gm * (r0[0] * r2[0] + r0[1] * r2[1] + r0[2] * r2[2]);
    tau[3] = gm * (r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]) + iso;
    tau[4] = gm * (r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2]);
    tau[5] = gm * (r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]) + iso;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const T k = c == 0 ? kx : ky;
        if (k != T(0)) {
            const T f[3] = { F.a[0][c], F.a[1][c], F.a[2][c] };
            const T n = dsqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
            const T cc = ddiv(k * (n - T(1)), n);
            tau[0] += cc * f[0] * f[0];
            tau[1] += cc * f[0] * f[1];
            tau[2] += cc * f[0] * f[2];
            tau[3] += cc * f[1] * f[1];
            tau[4] += cc * f[1] * f[2];
            tau[5] += cc * f[2] * f[2];
        }
    }
    *Jout = J;
}

// Neo-Hookean Kirchhoff stress tau = mu F F^T + (lambda log J - mu) I (NeoHookean.cpp:38-43).
template <class T>
__device__ __forceinline__ Mat3<T> neoHookeanTau(const Mat3<T>& F, T mu, T lambda)
{
    T J = det3(F);
    T scale = lambda * dlog(J) - mu;
    Mat3<T> tau = mulT(F, F);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) tau.a[i][j] = mu * tau.a[i][j] + (i == j ? scale : T(0));
    return tau;
}

// ------------------------------------------- FP32 form: H = F - I stored
// The float form keeps the deformation as H = F - I, and every stress below is
// written in terms of the strain-sized quantities H, J - 1 and E = F F^T - I,
// so no difference of two O(1) floats (det F - 1, F F^T - I, F - R) loses the
// strain to cancellation.  The models are the same functions of F as above.

// J - 1 = det(I + H) - 1 = tr H + (principal 2x2 minors of H) + det H.
template <class T>
__device__ __forceinline__ T jm1OfH(const Mat3<T>& H)
{
    const T h00 = H.a[0][0], h01 = H.a[0][1], h02 = H.a[0][2];
    const T h10 = H.a[1][0], h11 = H.a[1][1], h12 = H.a[1][2];
    const T h20 = H.a[2][0], h21 = H.a[2][1], h22 = H.a[2][2];
    const T m2 = (h00 * h11 - h01 * h10) + (h00 * h22 - h02 * h20) + (h11 * h22 - h12 * h21);
    return (h00 + h11 + h22) + m2 + det3(H);
}

// E = F F^T - I = H + H^T + H H^T, symmetric (xx xy xz yy yz zz).
template <class T>
__device__ __forceinline__ void strainOfH(const Mat3<T>& H, T* E)
{
    const T* r0 = H.a[0];
    const T* r1 = H.a[1];
    const T* r2 = H.a[2];
    E[0] = T(2) * r0[0] + (r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
    E[1] = (r0[1] + r1[0]) + (r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2]);
    E[2] = (r0[2] + r2[0]) + (r0[0] * r2[0] + r0[1] * r2[1] + r0[2] * r2[2]);
    E[3] = T(2) * r1[1] + (r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
    E[4] = (r1[2] + r2[1]) + (r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2]);
    E[5] = T(2) * r2[2] + (r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
}

template <class T>
__device__ __forceinline__ Mat3<T> minusIdentity(const Mat3<T>& F)
{
    Mat3<T> H = F;
#pragma unroll
    for (int i = 0; i < 3; i++) H.a[i][i] -= T(1);
    return H;
}

__device__ __forceinline__ Mat3<float> identityPlus(const Mat3<float>& H)
{
    Mat3<float> F = H;
#pragma unroll
    for (int i = 0; i < 3; i++) F.a[i][i] += 1.f;
    return F;
}

// Eigenvalues l1 >= l2 >= l3 of a symmetric 3x3 (xx xy xz yy yz zz) from the
// trigonometric solution of its characteristic cubic (O. K. Smith, "Eigenvalues
// of a symmetric 3x3 matrix", CACM 4, 1961), on the normalised deviator.
__device__ __forceinline__ void symEigvals3(const float* S, float& l1, float& l2, float& l3)
{
    const float q = (S[0] + S[3] + S[5]) * (1.f / 3.f);
    const float b0 = S[0] - q, b1 = S[3] - q, b2 = S[5] - q;
    const float p2 = b0 * b0 + b1 * b1 + b2 * b2 + 2.f * (S[1] * S[1] + S[2] * S[2] + S[4] * S[4]);
    if (!(p2 > 1e-30f)) {
        l1 = l2 = l3 = q;
        return;
    }
    const float p = __builtin_amdgcn_sqrtf(p2 * (1.f / 6.f));
    const float ip = __builtin_amdgcn_rcpf(p);
    const float c0 = b0 * ip, c1 = b1 * ip, c2 = b2 * ip, o01 = S[1] * ip, o02 = S[2] * ip, o12 = S[4] * ip;
    const float detB = c0 * (c1 * c2 - o12 * o12) - o01 * (o01 * c2 - o12 * o02) + o02 * (o01 * o12 - c1 * o02);
    const float r = fminf(1.f, fmaxf(-1.f, 0.5f * detB));
    const float phi = acosf(r) * (1.f / 3.f);
    l1 = q + 2.f * p * __cosf(phi);
    l3 = q + 2.f * p * __cosf(phi + 2.0943951023931953f);
    l2 = 3.f * q - l1 - l3;
}

// Fixed corotated Kirchhoff stress (the corotatedTau model above) from H.
// With B = F F^T = V^2 (left polar decomposition F = V R, J > 0), the reference's
// (F - R) F^T equals V^2 - V = f(B) for f(b) = b - sqrt(b), so
//   tau = 2 mu f(B) + lambda (J - 1) J I.
// f(B) is evaluated from the eigenvalues l_i of E = B - I (s_i = sqrt(1 + l_i)
// are the singular values of F) in Newton's divided-difference form, which
// stays exact for coincident eigenvalues and needs no eigenvectors:
//   f(B) = f1 I + f[1,2] (E - l1 I) + f[1,2,3] (E - l1 I)(E - l2 I),
//   f1 = s1 l1 / (s1 + 1),  f[1,2] = 1 - 1 / (s1 + s2),
//   f[1,2,3] = 1 / ((s1 + s2)(s2 + s3)(s1 + s3)).
// Below J = 1/20 (inverted or nearly, where the reference's signed SVD makes R
// differ from V^{-1} F) the reference formula runs on F = I + H.
// FAST: below J = 1/20 return true without a stress (the caller defers the
// particle to a kernel that runs the reference formula).
template <bool FAST = false>
__device__ __forceinline__ bool corotatedTauH(const Mat3<float>& H, float mu, float lambda, float* tau)
{
    const float jm1 = jm1OfH(H);
    const float J = 1.f + jm1;
    if (!(J > 0.05f)) {
        if constexpr (FAST) return true;
        const Mat3<float> t = corotatedTau(identityPlus(H), mu, lambda);
        tau[0] = t.a[0][0];
        tau[1] = 0.5f * (t.a[0][1] + t.a[1][0]);
        tau[2] = 0.5f * (t.a[0][2] + t.a[2][0]);
        tau[3] = t.a[1][1];
        tau[4] = 0.5f * (t.a[1][2] + t.a[2][1]);
        tau[5] = t.a[2][2];
        return false;
    }
    float E[6];
    strainOfH(H, E);
    float l1, l2, l3;
    symEigvals3(E, l1, l2, l3);
    const float s1 = __builtin_amdgcn_sqrtf(1.f + l1), s2 = __builtin_amdgcn_sqrtf(1.f + l2);
    const float s3 = __builtin_amdgcn_sqrtf(fmaxf(1.f + l3, 0.f));
    const float f1 = s1 * l1 * __builtin_amdgcn_rcpf(s1 + 1.f);
    const float d12 = 1.f - __builtin_amdgcn_rcpf(s1 + s2);
    const float d123 = __builtin_amdgcn_rcpf((s1 + s2) * (s2 + s3) * (s1 + s3));
    // M1 = E - l1 I, M2 = E - l2 I (they commute: M1 M2 is symmetric)
    const float a00 = E[0] - l1, a11 = E[3] - l1, a22 = E[5] - l1;
    const float b00 = E[0] - l2, b11 = E[3] - l2, b22 = E[5] - l2;
    const float e01 = E[1], e02 = E[2], e12 = E[4];
    const float p00 = a00 * b00 + e01 * e01 + e02 * e02;
    const float p01 = a00 * e01 + e01 * b11 + e02 * e12;
    const float p02 = a00 * e02 + e01 * e12 + e02 * b22;
    const float p11 = e01 * e01 + a11 * b11 + e12 * e12;
    const float p12 = e01 * e02 + a11 * e12 + e12 * b22;
    const float p22 = e02 * e02 + e12 * e12 + a22 * b22;
    const float two_mu = 2.f * mu;
    const float iso = lambda * jm1 * J;
    tau[0] = two_mu * (f1 + d12 * a00 + d123 * p00) + iso;
    tau[1] = two_mu * (d12 * e01 + d123 * p01);
    tau[2] = two_mu * (d12 * e02 + d123 * p02);
    tau[3] = two_mu * (f1 + d12 * a11 + d123 * p11) + iso;
    tau[4] = two_mu * (d12 * e12 + d123 * p12);
    tau[5] = two_mu * (f1 + d12 * a22 + d123 * p22) + iso;
    return false;
}

// QRStableNeoHookean Kirchhoff stress (qrStableTau above) from H:
// g mu F F^T + (vol - g mu J) I = g mu E + (J - 1)(g_l lambda J - g mu) I, and the
// fibre factor (|f| - 1) / |f| with |f| - 1 = (|f|^2 - 1) / (|f| + 1),
// |f_c|^2 - 1 = 2 H_cc + |H e_c|^2.
__device__ __forceinline__ void qrStableTauH(const Mat3<float>& H, float mu, float lambda, float g, float kx, float ky,
                                             float* tau, float* Jout)
{
    const float jm1 = jm1OfH(H);
    const float J = 1.f + jm1;
    const float gl = (jm1 >= 0.f) ? g * lambda : lambda;
    const float gm = g * mu;
    const float iso = jm1 * (gl * J - gm);
    float E[6];
    strainOfH(H, E);
    tau[0] = gm * E[0] + iso;
    tau[1] = gm * E[1];
    tau[2] = gm * E[2];
    tau[3] = gm * E[3] + iso;
    tau[4] = gm * E[4];
    tau[5] = gm * E[5] + iso;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const float k = c == 0 ? kx : ky;
        if (k != 0.f) {
            const float h[3] = { H.a[0][c], H.a[1][c], H.a[2][c] };
            const float f[3] = { h[0] + (c == 0 ? 1.f : 0.f), h[1] + (c == 1 ? 1.f : 0.f), h[2] };
            const float n2m1 = 2.f * h[c] + (h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
            const float n = __builtin_amdgcn_sqrtf(1.f + n2m1);
            const float cc = k * n2m1 * __builtin_amdgcn_rcpf((n + 1.f) * n);
            tau[0] += cc * f[0] * f[0];
            tau[1] += cc * f[0] * f[1];
            tau[2] += cc * f[0] * f[2];
            tau[3] += cc * f[1] * f[1];
            tau[4] += cc * f[1] * f[2];
            tau[5] += cc * f[2] * f[2];
        }
    }
    *Jout = J;
}

// QRAnisotropic Kirchhoff stress tau = P F^T from H = F - I, symmetric (xx xy xz
// yy yz zz), both forms.  The reference evaluates P = Q A R^{-T} from the Givens
// QR F = Q R (R_00, R_11 >= 0, R_22 with the sign of det F), A the symmetric
// matrix whose upper triangle is dPsiHat/dR R^T (QRAnisotropic.h:87-169,
// QrBasedHelper.h:28-40); so tau = Q A Q^T.  With dPsiHat/dR = diag(v / R_ii +
// c_i (R_ii - 1)) + g k_s (strict upper triangle of R), v = g_v k_v (J - 1) J,
// g_v = g for J >= 1 else 1, c_i = g k_i except c_0 = k_0 (undegraded first
// fibre) when not isotropic:
//   A = v I + g k_s (R R^T - diag(R_ii^2)) + diag(c_i (R_ii - 1) R_ii),
// whose entries are all strain-sized.  Q and R are Gram-Schmidt on the columns
// f_c of F (the Givens sweep's factors: q0 = f0 / |f0|, q1 the unit part of f1
// normal to q0, q2 = q0 x q1, R_22 = J / (R_00 R_11)), and R_ii - 1 come from H
// without cancellation:
//   R_00^2 - 1 = 2 H_00 + |h_0|^2,  R_11^2 - 1 = 2 H_11 + |h_1|^2 - R_01^2,
//   R_22 - 1 = ((J - 1) - (R_00 R_11 - 1)) / (R_00 R_11).
template <class T>
__device__ __forceinline__ void qrAnisoTauH(const Mat3<T>& H, T kv, T ks, T g, T k0, T k1, T k2, bool iso, T* tau,
                                            T* Jout)
{
    const T jm1 = jm1OfH(H);
    const T J = T(1) + jm1;
    T h[3][3], f[3][3]; // column c of H and of F
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            h[c][r] = H.a[r][c];
            f[c][r] = H.a[r][c] + (r == c ? T(1) : T(0));
        }
    const T n0m1 = T(2) * h[0][0] + (h[0][0] * h[0][0] + h[0][1] * h[0][1] + h[0][2] * h[0][2]);
    const T R00 = dsqrt(T(1) + n0m1);
    const T r00m1 = ddiv(n0m1, R00 + T(1));
    const T i00 = ddiv(T(1), R00);
    const T q0[3] = { f[0][0] * i00, f[0][1] * i00, f[0][2] * i00 };
    const T R01 = q0[0] * f[1][0] + q0[1] * f[1][1] + q0[2] * f[1][2];
    const T R02 = q0[0] * f[2][0] + q0[1] * f[2][1] + q0[2] * f[2][2];
    const T n1m1 = T(2) * h[1][1] + (h[1][0] * h[1][0] + h[1][1] * h[1][1] + h[1][2] * h[1][2]) - R01 * R01;
    const T R11 = dsqrt(T(1) + n1m1);
    const T r11m1 = ddiv(n1m1, R11 + T(1));
    const T i11 = ddiv(T(1), R11);
    const T q1[3] = { (f[1][0] - R01 * q0[0]) * i11, (f[1][1] - R01 * q0[1]) * i11, (f[1][2] - R01 * q0[2]) * i11 };
    const T q2[3] = { q0[1] * q1[2] - q0[2] * q1[1], q0[2] * q1[0] - q0[0] * q1[2], q0[0] * q1[1] - q0[1] * q1[0] };
    const T R12 = q1[0] * f[2][0] + q1[1] * f[2][1] + q1[2] * f[2][2];
    const T pr = R00 * R11;
    const T prm1 = r00m1 + r11m1 + r00m1 * r11m1;
    const T ipr = ddiv(T(1), pr);
    const T R22 = J * ipr;
    const T r22m1 = (jm1 - prm1) * ipr;
    const T v = ((jm1 >= T(0)) ? g : T(1)) * kv * jm1 * J;
    const T gks = g * ks;
    const T c0 = iso ? g * k0 : k0;
    T A[3][3];
    A[0][0] = v + gks * (R01 * R01 + R02 * R02) + c0 * r00m1 * R00;
    A[1][1] = v + gks * (R12 * R12) + g * k1 * r11m1 * R11;
    A[2][2] = v + g * k2 * r22m1 * R22;
    A[0][1] = A[1][0] = gks * (R01 * R11 + R02 * R12);
    A[0][2] = A[2][0] = gks * (R02 * R22);
    A[1][2] = A[2][1] = gks * (R12 * R22);
    const T* q[3] = { q0, q1, q2 };
    // B = Q A (row a: sum_i q_i[a] A_ij), tau = B Q^T
    T B[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int j = 0; j < 3; j++) B[a][j] = q[0][a] * A[0][j] + q[1][a] * A[1][j] + q[2][a] * A[2][j];
    const int ia[6] = { 0, 0, 0, 1, 1, 2 }, ib[6] = { 0, 1, 2, 1, 2, 2 };
#pragma unroll
    for (int e = 0; e < 6; e++) {
        const int a = ia[e], b = ib[e];
        tau[e] = B[a][0] * q[0][b] + B[a][1] * q[1][b] + B[a][2] * q[2][b];
    }
    *Jout = J;
}

// J^{-2/3} in double, out of line: inlined, the double pow raises the register
// count of every kernel that holds the NeoHookeanBorden branch.
__device__ __noinline__ double powMinusTwoThirds(double J) { return pow(J, -2.0 / 3.0); }

// NeoHookeanBorden Kirchhoff stress from H = F - I, symmetric, both forms
// (NeoHookeanBorden.cpp:78-92): mu J^{-2/3} dev(F F^T) + J kappa/2 (J - 1/J) I,
// the deviatoric part times g and the volumetric part times g only for J >= 1.
// dev(F F^T) = dev(E) and J kappa/2 (J - 1/J) = kappa/2 (J - 1)(J + 1) are
// strain-sized.
template <class T>
__device__ __forceinline__ void bordenTauH(const Mat3<T>& H, T mu, T kappa, T g, T* tau, T* Jout)
{
    const T jm1 = jm1OfH(H);
    const T J = T(1) + jm1;
    T E[6];
    strainOfH(H, E);
    T jpow; // J^{-2/3}
    if constexpr (sizeof(T) == 4)
        jpow = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(J) * (-2.f / 3.f));
    else
        jpow = powMinusTwoThirds(J);
    const T sd = g * mu * jpow;
    const T tr3 = (E[0] + E[3] + E[5]) * (T(1) / T(3));
    const T vol = ((jm1 >= T(0)) ? g : T(1)) * (kappa * T(0.5)) * jm1 * (T(2) + jm1);
    tau[0] = sd * (E[0] - tr3) + vol;
    tau[1] = sd * E[1];
    tau[2] = sd * E[2];
    tau[3] = sd * (E[3] - tr3) + vol;
    tau[4] = sd * E[4];
    tau[5] = sd * (E[5] - tr3) + vol;
    *Jout = J;
}

// Neo-Hookean Kirchhoff stress (neoHookeanTau above) from H:
// mu F F^T + (lambda log J - mu) I = mu E + lambda log1p(J - 1) I.
__device__ __forceinline__ void neoHookeanTauH(const Mat3<float>& H, float mu, float lambda, float* tau)
{
    const float lj = lambda * log1pf(jm1OfH(H));
    float E[6];
    strainOfH(H, E);
    tau[0] = mu * E[0] + lj;
    tau[1] = mu * E[1];
    tau[2] = mu * E[2];
    tau[3] = mu * E[3] + lj;
    tau[4] = mu * E[4];
    tau[5] = mu * E[5] + lj;
}

// ------------------------------------------------------ symmetric eigen (Jacobi)
// Unit eigenvectors in the columns of Q.  Stands in for Eigen::EigenSolver at
// AnisoFractureSimulation.h:244-254 (sigma^+ does not depend on the basis
// chosen in a degenerate eigenspace or on eigenvector signs).
template <class T>
__device__ __forceinline__ void jacobiRot(Mat3<T>& S, Mat3<T>& Q, int p, int q)
{
    T apq = S.a[p][q];
    if (apq == T(0)) return;
    T theta = (S.a[q][q] - S.a[p][p]) / (T(2) * apq);
    T t = (theta >= T(0) ? T(1) : T(-1)) / (dabs(theta) + dsqrt(theta * theta + T(1)));
    T c = T(1) / dsqrt(t * t + T(1)), s = t * c;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        T skp = S.a[k][p], skq = S.a[k][q];
        S.a[k][p] = c * skp - s * skq;
        S.a[k][q] = s * skp + c * skq;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        T spk = S.a[p][k], sqk = S.a[q][k];
        S.a[p][k] = c * spk - s * sqk;
        S.a[q][k] = s * spk + c * sqk;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        T qkp = Q.a[k][p], qkq = Q.a[k][q];
        Q.a[k][p] = c * qkp - s * qkq;
        Q.a[k][q] = s * qkp + c * qkq;
    }
}

template <class T>
__device__ __forceinline__ void symEigen(Mat3<T> S, T* lam, Mat3<T>& Q)
{
    setIdentity(Q);
    constexpr int max_sweeps = sizeof(T) == 4 ? 6 : 10;
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        T off = S.a[0][1] * S.a[0][1] + S.a[0][2] * S.a[0][2] + S.a[1][2] * S.a[1][2];
        T diag = S.a[0][0] * S.a[0][0] + S.a[1][1] * S.a[1][1] + S.a[2][2] * S.a[2][2];
        // off-diagonal mass down to rounding level: (4 eps)^2 of the diagonal
        if (off <= T(16) * Lim<T>::eps * Lim<T>::eps * diag || off == T(0)) break;
        jacobiRot(S, Q, 0, 1);
        jacobiRot(S, Q, 0, 2);
        jacobiRot(S, Q, 1, 2);
    }
    lam[0] = S.a[0][0];
    lam[1] = S.a[1][1];
    lam[2] = S.a[2][2];
}

// sigma^+ = sum_i <lambda_i>_+ v_i v_i^T of a symmetric 3x3.  Generic form: Jacobi.
template <class T>
__device__ __forceinline__ Mat3<T> positivePart(const Mat3<T>& S)
{
    T lam[3];
    Mat3<T> Qe, sp;
    symEigen(S, lam, Qe);
#pragma unroll
    for (int e = 0; e < 3; e++) lam[e] = (lam[e] + dabs(lam[e])) / T(2);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            sp.a[i][j] = lam[0] * (Qe.a[i][0] * Qe.a[j][0]) + lam[1] * (Qe.a[i][1] * Qe.a[j][1]) + lam[2] * (Qe.a[i][2] * Qe.a[j][2]);
    return sp;
}

// Unit eigenvector of the symmetric matrix with diagonal d and off-diagonal
// (o01, o02, o12) for its simple eigenvalue l: the largest cross product of two
// rows of M - l I.  Called on the normalised deviator (entries O(1)), so the
// squared norms stay far from the float underflow range; a vanishing cross
// product (l degenerate to rounding) gives v = 0.
__device__ __forceinline__ void eigvecOf(const float* dg, float o01, float o02, float o12, float l, float* v)
{
    const float r0[3] = { dg[0] - l, o01, o02 };
    const float r1[3] = { o01, dg[1] - l, o12 };
    const float r2[3] = { o02, o12, dg[2] - l };
    float c[3][3];
    const float* A[3] = { r0, r0, r1 };
    const float* Bv[3] = { r1, r2, r2 };
    float best = -1.f;
    int bi = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        c[k][0] = A[k][1] * Bv[k][2] - A[k][2] * Bv[k][1];
        c[k][1] = A[k][2] * Bv[k][0] - A[k][0] * Bv[k][2];
        c[k][2] = A[k][0] * Bv[k][1] - A[k][1] * Bv[k][0];
        const float n = c[k][0] * c[k][0] + c[k][1] * c[k][1] + c[k][2] * c[k][2];
        if (n > best) {
            best = n;
            bi = k;
        }
    }
    const float inv = best > 1e-30f ? __builtin_amdgcn_rsqf(best) : 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = (bi == 0 ? c[0][a] : (bi == 1 ? c[1][a] : c[2][a])) * inv;
}

// Float form: eigenvalues from the trigonometric solution of the characteristic
// cubic (O. K. Smith, "Eigenvalues of a symmetric 3x3 matrix", CACM 4, 1961) of
// the normalised deviator B = (S - q I) / p, and at most one eigenvector (of B,
// the same as S's) -- of the eigenvalue on the other side of zero from the
// rest -- so no iteration.  When that eigenvalue is nearly degenerate with a
// neighbour it is itself near zero and its term in sigma^+ is negligible.  A
// deviator below 1e-6 of the mean stress (or 1e-15 absolute) counts as
// isotropic: every eigenvalue then has the sign of q.
template <>
__device__ __forceinline__ Mat3<float> positivePart<float>(const Mat3<float>& S)
{
    const float a01 = S.a[0][1], a02 = S.a[0][2], a12 = S.a[1][2];
    const float q = (S.a[0][0] + S.a[1][1] + S.a[2][2]) * (1.f / 3.f);
    const float b00 = S.a[0][0] - q, b11 = S.a[1][1] - q, b22 = S.a[2][2] - q;
    const float p1 = a01 * a01 + a02 * a02 + a12 * a12;
    const float p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.f * p1;
    Mat3<float> out;
    if (!(p2 > fmaxf(1e-30f, 1e-12f * q * q))) { // isotropic
        if (q >= 0.f) return S;
        setIdentity(out);
#pragma unroll
        for (int i = 0; i < 3; i++) out.a[i][i] = 0.f;
        return out;
    }
    const float p = __builtin_amdgcn_sqrtf(p2 * (1.f / 6.f));
    const float ip = __builtin_amdgcn_rcpf(p);
    const float dg[3] = { b00 * ip, b11 * ip, b22 * ip };
    const float c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
    const float detB = dg[0] * (dg[1] * dg[2] - c12 * c12) - c01 * (c01 * dg[2] - c12 * c02) + c02 * (c01 * c12 - dg[1] * c02);
    const float r = fminf(1.f, fmaxf(-1.f, 0.5f * detB));
    const float phi = acosf(r) * (1.f / 3.f);
    const float e1 = 2.f * __cosf(phi), e3 = 2.f * __cosf(phi + 2.0943951023931953f);
    const float l1 = q + p * e1, l3 = q + p * e3;
    const float l2 = 3.f * q - l1 - l3;
    if (l3 >= 0.f) return S;
    if (l1 <= 0.f) {
        setIdentity(out);
#pragma unroll
        for (int i = 0; i < 3; i++) out.a[i][i] = 0.f;
        return out;
    }
    float v[3];
    if (l2 <= 0.f) { // only l1 > 0
        eigvecOf(dg, c01, c02, c12, e1, v);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) out.a[i][j] = l1 * (v[i] * v[j]);
    }
    else { // l1, l2 > 0 >= l3
        eigvecOf(dg, c01, c02, c12, e3, v);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) out.a[i][j] = S.a[i][j] - l3 * (v[i] * v[j]);
    }
    return out;
}

template <class T>
__device__ __forceinline__ T macaulay(T x) { return (x + dabs(x)) / T(2); } // MathTools.h:276-281

// ------------------------------------------------------ quadratic B-spline
// Weights of computeBSplineWeights (BSplines.h:72-98) from the offset
// d0 = x_index - base in [0.5, 1.5) (the base node is cellOf's, zmpm_device.hpp).
template <class T>
__device__ __forceinline__ void bsplineFrac(T d0, T* w, T* dw)
{
    T z = T(1.5) - d0;
    w[0] = T(0.5) * z * z;
    T d1 = d0 - T(1);
    w[1] = T(0.75) - d1 * d1;
    T d2 = T(1) - d1;
    T zz = T(1.5) - d2;
    w[2] = T(0.5) * zz * zz;
    dw[0] = -z;
    dw[1] = T(-2) * d1;
    dw[2] = zz;
}

// ------------------------------------------------------ SPGrid bit layout
// Bit_Spread_u64 (SPGrid_Utilities.h:81-131) for non-negative coordinates:
// deposit the low bits of v into the set bits of mask, low to high.
__device__ __forceinline__ uint64_t bitSpread(uint32_t v, uint64_t mask)
{
    uint64_t r = 0;
    uint32_t bit = 0;
    while (mask && bit < 13) {
        uint64_t low = mask & (~mask + 1);
        if (v & (1u << bit)) r |= low;
        mask ^= low;
        bit++;
    }
    return r;
}

} // namespace zmpm
