// geometry.cpp -- boundary conditions and the initial curve.
#include <algorithm>
#include <array>
#include <cmath>

#include "twisty_host.hpp"

namespace twisty {

namespace PerturbUtils {
std::array<float, 13> ToCudaSafe(const BoundaryConditions &bc)
{
    return { bc.m_startPos.x, bc.m_startPos.y, bc.m_startPos.z, bc.m_startDir.x, bc.m_startDir.y, bc.m_startDir.z,
        bc.m_endPos.x, bc.m_endPos.y, bc.m_endPos.z, bc.m_endDir.x, bc.m_endDir.y, bc.m_endDir.z, bc.arclength };
}
}  // namespace PerturbUtils

namespace PathGeneration {

// PathGeneration.cpp:12-43: the arclength at which the N-2 free chords between
// the hinges are collinear, |Xe - Xs - (L/M)(Ns + Ne)| = (M - 2) L / M, as a
// quadratic in L (float arithmetic as the reference).
float CalculateMinimumArclength(const PerturbUtils::BoundaryConditions &bc, uint32_t numSegmentsPerCurve)
{
    const float M = (float)numSegmentsPerCurve;
    const Vec3 S = bc.m_startDir + bc.m_endDir;
    const Vec3 D = bc.m_endPos - bc.m_startPos;
    const float a = S.Dot(S) - ((M - 4.0f) * M + 4.0f);
    const float b = -2.0f * M * S.Dot(D);
    const float c = M * M * D.Dot(D);
    const float r1 = (-b - std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    const float r2 = (-b + std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    float m = 1000000.0f;
    if (!std::isnan(r1) && r1 > 0.0f) m = std::min(m, r1);
    if (!std::isnan(r2) && r2 > 0.0f) m = std::min(m, r2);
    return m;
}

// Hinges as Curve::SetBoundaryConditions (Curve.cpp:20-30); the M = N-2 free
// chords of length ds form a planar circular arc from p1 to p_{N-1} that bulges
// toward startDir's component across the chord (the least-aligned coordinate
// axis when that component vanishes).  With half-opening angle Phi and radius
// R: |p_{N-1} - p1| = 2 R sin(Phi), ds = 2 R sin(Phi / M); Phi in [0, pi) is
// found by bisection on sin(Phi) / sin(Phi / M), which falls from M to 0.
namespace {
// bulgeIn (nullable) replaces startDir as the source of the bulge direction;
// bulgeOut (nullable) receives the unit bulge direction used; *degenerate is set
// when startDir lies on the chord and the least-aligned axis was taken instead.
bool arc_f64(const std::array<double, 3> &s0, const std::array<double, 3> &sd, const std::array<double, 3> &e0,
      const std::array<double, 3> &ed, float arclength, uint32_t N, std::vector<std::array<double, 3>> &P,
      const double *bulgeIn = nullptr, double *bulgeOut = nullptr, bool *degenerate = nullptr);
}

bool InitialArc(const PerturbUtils::BoundaryConditions &bc, uint32_t N, std::vector<Vec3> &positions)
{
    std::vector<double> P;
    if (!InitialArcF64(bc, N, P)) return false;
    positions.assign(N + 1, Vec3());
    for (uint32_t i = 0; i <= N; ++i)
        positions[i] = Vec3((float)P[3 * i], (float)P[3 * i + 1], (float)P[3 * i + 2]);
    return true;
}

bool InitialArcF64(const PerturbUtils::BoundaryConditions &bc, uint32_t N, std::vector<double> &positions)
{
    if (N < 3 || !(bc.arclength > 0.0f)) return false;
    auto d3 = [](const Vec3 &v) { return std::array<double, 3> { v.x, v.y, v.z }; };
    std::vector<std::array<double, 3>> P;
    if (!arc_f64(d3(bc.m_startPos), d3(bc.m_startDir), d3(bc.m_endPos), d3(bc.m_endDir), bc.arclength, N, P))
        return false;
    positions.resize(3 * (N + 1));
    for (uint32_t i = 0; i <= N; ++i)
        for (int a = 0; a < 3; ++a) positions[3 * i + a] = P[i][a];
    return true;
}

bool InitialArcExitJacobian(const PerturbUtils::BoundaryConditions &bc, uint32_t N, std::vector<double> &dTheta,
      std::vector<double> &dPhi)
{
    if (N < 3 || !(bc.arclength > 0.0f)) return false;
    auto d3 = [](const Vec3 &v) { return std::array<double, 3> { v.x, v.y, v.z }; };
    const auto s0 = d3(bc.m_startPos), sd = d3(bc.m_startDir), e0 = d3(bc.m_endPos);
    const double ex = bc.m_endDir.x, ey = bc.m_endDir.y, ez = bc.m_endDir.z;
    const double en = std::sqrt(ex * ex + ey * ey + ez * ez);
    const double theta = std::acos(std::max(-1.0, std::min(1.0, ez / en)));
    const double phi = std::atan2(ey, ex);
    // When startDir lies on the chord the bulge plane is a free choice (the
    // least-aligned axis) and flips under any tilt of the end direction, so the
    // curve is not differentiable there; the derivative is then taken with the
    // bulge direction of the base curve held fixed.
    const std::array<double, 3> ed0 { std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta) };
    std::vector<std::array<double, 3>> Pp, Pm;
    double bulge[3];
    bool degenerate = false;
    if (!arc_f64(s0, sd, e0, ed0, bc.arclength, N, Pp, nullptr, bulge, &degenerate)) return false;
    auto curve = [&](double th, double ph, std::vector<std::array<double, 3>> &P) {
        const std::array<double, 3> ed { std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th) };
        return arc_f64(s0, sd, e0, ed, bc.arclength, N, P, degenerate ? bulge : nullptr);
    };
    auto central = [&](double dth, double dph, double h, std::vector<double> &out) {
        if (!curve(theta + dth * h, phi + dph * h, Pp) || !curve(theta - dth * h, phi - dph * h, Pm)) return false;
        out.resize(3 * (N + 1));
        for (uint32_t i = 0; i <= N; ++i)
            for (int a = 0; a < 3; ++a) out[3 * i + a] = (Pp[i][a] - Pm[i][a]) / (2.0 * h);
        return true;
    };
    const double h = 1e-4;
    for (int which = 0; which < 2; ++which) {
        std::vector<double> d1, d2;
        const double dth = which == 0 ? 1.0 : 0.0, dph = which == 0 ? 0.0 : 1.0;
        if (!central(dth, dph, h, d1) || !central(dth, dph, 0.5 * h, d2)) return false;
        std::vector<double> &out = which == 0 ? dTheta : dPhi;
        out.resize(d1.size());
        for (size_t k = 0; k < d1.size(); ++k) out[k] = (4.0 * d2[k] - d1[k]) / 3.0;
    }
    return true;
}

namespace {
bool arc_f64(const std::array<double, 3> &s0, const std::array<double, 3> &sd, const std::array<double, 3> &e0,
      const std::array<double, 3> &ed, float arclength, uint32_t N, std::vector<std::array<double, 3>> &P,
      const double *bulgeIn, double *bulgeOut, bool *degenerate)
{
    const double ds = (double)(arclength / (float)N);
    P.assign(N + 1, { 0.0, 0.0, 0.0 });
    for (int a = 0; a < 3; ++a) {
        P[0][a] = s0[a];
        P[1][a] = s0[a] + ds * sd[a];
        P[N - 1][a] = e0[a] - ds * ed[a];
        P[N][a] = e0[a];
    }
    const uint32_t M = N - 2;
    double u[3], c2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        u[a] = P[N - 1][a] - P[1][a];
        c2 += u[a] * u[a];
    }
    const double c = std::sqrt(c2);
    if (!(c > 0.0) || c > (double)M * ds * (1.0 + 1e-12)) return false;
    for (double &x : u) x /= c;
    // bulge direction v: startDir (or the given bulge) minus its chord component
    const double *bs = bulgeIn ? bulgeIn : sd.data();
    double v[3], vd = bs[0] * u[0] + bs[1] * u[1] + bs[2] * u[2], vn = 0.0;
    for (int a = 0; a < 3; ++a) {
        v[a] = bs[a] - vd * u[a];
        vn += v[a] * v[a];
    }
    if (degenerate) *degenerate = std::sqrt(vn) < 1e-9;
    if (std::sqrt(vn) < 1e-9) {
        int k = 0;
        if (std::fabs(u[1]) < std::fabs(u[0])) k = 1;
        if (std::fabs(u[2]) < std::min(std::fabs(u[0]), std::fabs(u[1]))) k = 2;
        double ax[3] = { 0, 0, 0 };
        ax[k] = 1.0;
        const double d = ax[0] * u[0] + ax[1] * u[1] + ax[2] * u[2];
        vn = 0.0;
        for (int a = 0; a < 3; ++a) {
            v[a] = ax[a] - d * u[a];
            vn += v[a] * v[a];
        }
    }
    vn = std::sqrt(vn);
    for (double &x : v) x /= vn;
    if (bulgeOut)
        for (int a = 0; a < 3; ++a) bulgeOut[a] = v[a];
    const double ratio = c / ds;
    if (ratio >= (double)M * (1.0 - 1e-15)) {
        for (uint32_t j = 1; j < M; ++j)
            for (int a = 0; a < 3; ++a) P[1 + j][a] = P[1][a] + ds * (double)j * u[a];
    } else {
        double lo = 0.0, hi = M_PI;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (std::sin(mid) / std::sin(mid / (double)M) > ratio) lo = mid;
            else hi = mid;
        }
        const double Phi = 0.5 * (lo + hi);
        const double alpha = 2.0 * Phi / (double)M;  // turning angle per chord
        const double R = ds / (2.0 * std::sin(0.5 * alpha));
        double C[3];
        for (int a = 0; a < 3; ++a) C[a] = 0.5 * (P[1][a] + P[N - 1][a]) - R * std::cos(Phi) * v[a];
        for (uint32_t j = 1; j < M; ++j) {
            const double phi = -Phi + (double)j * alpha;
            for (int a = 0; a < 3; ++a) P[1 + j][a] = C[a] + R * (std::sin(phi) * u[a] + std::cos(phi) * v[a]);
        }
    }
    return true;
}
}  // namespace

}  // namespace PathGeneration
}  // namespace twisty
