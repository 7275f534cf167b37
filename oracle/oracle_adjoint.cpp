// oracle_adjoint.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 reverse-mode derivative of the log-estimator E = log10(sum_p w_p / P)
// over K-perturbation paths that restart from one initial curve, written from
// the mathematics, to check the library's adjoint kernel.  The forward here is
// the oracle's path (twisty_oracle.h: draws, sub-chain rotation of
// FullExperimentRunnerOptimalPerturb.cpp:495-519, tangents/curvatures of
// CurvePerturbUtils.cpp:147-166,219-243, weight of PathWeighters.cpp:27-104) in
// segment form: the rotation of points left+1..right-1 about p_left is the
// rotation of the segment vectors d_left..d_{right-1}, and the axis is their sum.
// Every intermediate state is kept, so the reverse sweep needs no rewinding.
//
// lerpMode 0 is the reference's weight lookup (fraction in curvature units,
// float arithmetic); lerpMode 1 is a continuous variant (fraction in bin units,
// FP64) whose derivative in the curvature is exactly the bin-average slope
// (T[r] - T[l]) / step that the adjoint uses in both modes.  Mode 1 exists so a
// central difference can pin the adjoint formula: in mode 0 the weight jumps
// at every bin edge and a finite difference sees the jumps.
#include "twisty_oracle.h"

#include <cmath>
#include <thread>
#include <vector>

using namespace twisty_oracle;

namespace {

struct AdjTable {
    const float *table;
    int32_t size;
    float minK, maxK, step, absorption, ds;
};

struct Lerp {
    int32_t l, r;
    double frac;    // coefficient of T[r] (1 - frac multiplies T[l])
    double value;
    double slope;   // d value / d kappa used by the adjoint
};

Lerp LerpAt(const AdjTable &t, double kappa, int mode)
{
    Lerp o;
    if (mode == 0) {
        // PathWeighters.cpp:41-82 on the float curvature.
        float curvature = (float)kappa;
        bool clamped = false;
        if (curvature < t.minK) {
            curvature = t.minK;
            clamped = true;
        }
        if (curvature > t.maxK) {
            curvature = t.maxK;
            clamped = true;
        }
        const float distanceFromMin = curvature - t.minK;
        const float realIdx = distanceFromMin / t.step;
        int32_t leftIdx = (int32_t)std::floor(realIdx);
        if (leftIdx > t.size - 1) leftIdx = t.size - 1;
        if (leftIdx < 0) leftIdx = 0;
        const int32_t rightIdx = (leftIdx == t.size - 1) ? leftIdx : leftIdx + 1;
        const float interpDist = distanceFromMin - ((float)leftIdx * t.step);
        volatile float a = t.table[leftIdx] * (1.0f - interpDist);
        volatile float b = t.table[rightIdx] * interpDist;
        o.l = leftIdx;
        o.r = rightIdx;
        o.frac = interpDist;
        o.value = (double)(float)(a + b);
        o.slope = clamped ? 0.0 : ((double)t.table[rightIdx] - (double)t.table[leftIdx]) / (double)t.step;
    } else {
        double k = kappa;
        bool clamped = false;
        if (k < t.minK) {
            k = t.minK;
            clamped = true;
        }
        if (k > t.maxK) {
            k = t.maxK;
            clamped = true;
        }
        const double u = (k - (double)t.minK) / (double)t.step;
        int32_t leftIdx = (int32_t)std::floor(u);
        if (leftIdx > t.size - 1) leftIdx = t.size - 1;
        if (leftIdx < 0) leftIdx = 0;
        const int32_t rightIdx = (leftIdx == t.size - 1) ? leftIdx : leftIdx + 1;
        const double f = u - leftIdx;
        o.l = leftIdx;
        o.r = rightIdx;
        o.frac = f;
        o.value = (double)t.table[leftIdx] * (1.0 - f) + (double)t.table[rightIdx] * f;
        o.slope = clamped ? 0.0 : ((double)t.table[rightIdx] - (double)t.table[leftIdx]) / (double)t.step;
    }
    return o;
}

struct Step {
    uint32_t left, right;
    double n[3], invA;
    double c, s, C;
    std::vector<V3<double>> pre;  // segments left..right-1 before the rotation
};

struct PathWork {
    std::vector<V3<double>> d, t, lam;
    std::vector<double> kappa;
    std::vector<Lerp> lerp;
    std::vector<Step> steps;
};

inline void RotMat(const double n[3], double c, double s, double C, double R[3][3])
{
    R[0][0] = c + C * n[0] * n[0];
    R[0][1] = C * n[0] * n[1] - s * n[2];
    R[0][2] = C * n[0] * n[2] + s * n[1];
    R[1][0] = C * n[1] * n[0] + s * n[2];
    R[1][1] = c + C * n[1] * n[1];
    R[1][2] = C * n[1] * n[2] - s * n[0];
    R[2][0] = C * n[2] * n[0] - s * n[1];
    R[2][1] = C * n[2] * n[1] + s * n[0];
    R[2][2] = c + C * n[2] * n[2];
}

// Forward of one path; returns ln w without absorption (or NaN when invalid).
double ForwardPath(const std::vector<V3<double>> &init, uint32_t N, double invDs, uint32_t K, uint64_t seed,
      uint64_t path, const AdjTable &t, int mode, PathWork &w)
{
    w.d.resize(N);
    for (uint32_t i = 0; i < N; ++i)
        w.d[i] = { init[i + 1].x - init[i].x, init[i + 1].y - init[i].y, init[i + 1].z - init[i].z };
    w.steps.resize(K);
    for (uint32_t k = 0; k < K; ++k) {
        const Draw dr = draw_perturbation(seed, 0, path, k, N);
        Step &st = w.steps[k];
        st.left = dr.left;
        st.right = dr.left + dr.diff;
        double a[3] = { 0, 0, 0 };
        for (uint32_t i = st.left; i < st.right; ++i) {
            a[0] += w.d[i].x;
            a[1] += w.d[i].y;
            a[2] += w.d[i].z;
        }
        const double an = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        st.invA = 1.0 / an;
        for (int c = 0; c < 3; ++c) st.n[c] = a[c] * st.invA;
        const double half = kPi * (double)dr.x;  // rotation angle 2 pi x
        st.c = std::cos(2.0 * half);
        st.s = std::sin(2.0 * half);
        st.C = 2.0 * std::sin(half) * std::sin(half);
        double R[3][3];
        RotMat(st.n, st.c, st.s, st.C, R);
        st.pre.assign(w.d.begin() + st.left, w.d.begin() + st.right);
        for (uint32_t i = st.left; i < st.right; ++i) {
            const V3<double> v = w.d[i];
            w.d[i] = { R[0][0] * v.x + R[0][1] * v.y + R[0][2] * v.z, R[1][0] * v.x + R[1][1] * v.y + R[1][2] * v.z,
                R[2][0] * v.x + R[2][1] * v.y + R[2][2] * v.z };
        }
    }
    w.t.resize(N);
    for (uint32_t i = 0; i < N; ++i) {
        const double inv = 1.0 / std::sqrt(w.d[i].x * w.d[i].x + w.d[i].y * w.d[i].y + w.d[i].z * w.d[i].z);
        w.t[i] = { w.d[i].x * inv, w.d[i].y * inv, w.d[i].z * inv };
    }
    w.kappa.resize(N - 1);
    w.lerp.resize(N - 1);
    double lnw = 0.0;
    for (uint32_t q = 0; q + 1 < N; ++q) {
        const double e0 = w.t[q + 1].x - w.t[q].x, e1 = w.t[q + 1].y - w.t[q].y, e2 = w.t[q + 1].z - w.t[q].z;
        w.kappa[q] = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2) * invDs;
        w.lerp[q] = LerpAt(t, w.kappa[q], mode);
        lnw += std::log(w.lerp[q].value);
    }
    return lnw;
}

// Adds scale * d ln w / d(table) to dT and scale * d ln w / d(initial segments)
// to dSeg (when non-null) for the path in w.
void AdjointPath(uint32_t N, double invDs, PathWork &w, const AdjTable &t, double scale, double *dT, double *dSeg)
{
    std::vector<V3<double>> v(N - 1);
    for (uint32_t q = 0; q + 1 < N; ++q) {
        const Lerp &L = w.lerp[q];
        const double cw = scale / L.value;
        dT[L.l] += cw * (1.0 - L.frac);
        dT[L.r] += cw * L.frac;
        const double e0 = (w.t[q + 1].x - w.t[q].x) * invDs, e1 = (w.t[q + 1].y - w.t[q].y) * invDs,
                     e2 = (w.t[q + 1].z - w.t[q].z) * invDs;
        const double kap = w.kappa[q];
        const double g = kap > 0.0 ? L.slope / L.value * invDs / kap : 0.0;
        v[q] = { g * e0, g * e1, g * e2 };
    }
    if (!dSeg) return;
    w.lam.resize(N);
    for (uint32_t i = 0; i < N; ++i) {
        double m[3] = { 0, 0, 0 };
        if (i >= 1) {
            m[0] += v[i - 1].x;
            m[1] += v[i - 1].y;
            m[2] += v[i - 1].z;
        }
        if (i + 1 < N) {
            m[0] -= v[i].x;
            m[1] -= v[i].y;
            m[2] -= v[i].z;
        }
        const V3<double> &ti = w.t[i];
        const double tm = ti.x * m[0] + ti.y * m[1] + ti.z * m[2];
        const double inv = 1.0 / std::sqrt(w.d[i].x * w.d[i].x + w.d[i].y * w.d[i].y + w.d[i].z * w.d[i].z);
        w.lam[i] = { (m[0] - tm * ti.x) * inv * scale, (m[1] - tm * ti.y) * inv * scale,
            (m[2] - tm * ti.z) * inv * scale };
    }
    for (uint32_t k = (uint32_t)w.steps.size(); k-- > 0;) {
        const Step &st = w.steps[k];
        double R[3][3];
        RotMat(st.n, st.c, st.s, st.C, R);
        double G[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
        for (uint32_t i = st.left; i < st.right; ++i) {
            const V3<double> &p = st.pre[i - st.left];
            const V3<double> &l = w.lam[i];
            const double lv[3] = { l.x, l.y, l.z }, pv[3] = { p.x, p.y, p.z };
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) G[a][b] += lv[a] * pv[b];
        }
        const double *n = st.n;
        double h[3];
        h[0] = st.s * (G[2][1] - G[1][2]);
        h[1] = st.s * (G[0][2] - G[2][0]);
        h[2] = st.s * (G[1][0] - G[0][1]);
        for (int kk = 0; kk < 3; ++kk) {
            double gn = 0.0, gtn = 0.0;
            for (int b = 0; b < 3; ++b) {
                gn += G[kk][b] * n[b];
                gtn += G[b][kk] * n[b];
            }
            h[kk] += st.C * (gn + gtn);
        }
        const double hn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
        const double da[3] = { (h[0] - hn * n[0]) * st.invA, (h[1] - hn * n[1]) * st.invA,
            (h[2] - hn * n[2]) * st.invA };
        for (uint32_t i = st.left; i < st.right; ++i) {
            const V3<double> l = w.lam[i];
            w.lam[i] = { R[0][0] * l.x + R[1][0] * l.y + R[2][0] * l.z + da[0],
                R[0][1] * l.x + R[1][1] * l.y + R[2][1] * l.z + da[1],
                R[0][2] * l.x + R[1][2] * l.y + R[2][2] * l.z + da[2] };
        }
    }
    for (uint32_t i = 0; i < N; ++i) {
        dSeg[3 * i + 0] += w.lam[i].x;
        dSeg[3 * i + 1] += w.lam[i].y;
        dSeg[3 * i + 2] += w.lam[i].z;
    }
}

}  // namespace

extern "C" {

// FP64 estimator and its adjoint for paths [0, nPaths) (seed, stream 0).
//   initialPos: (N+1)*3 doubles.  lerpMode: 0 reference lookup, 1 continuous.
//   log10SumW: log10 sum_valid w (with absorption, as oracle_simulate).
//   dTable (tableSize) and dSeg (3N): dE/dT and dE/d(initial segments) with
//   E = log10(sum w / nPaths); either may be NULL (both NULL = estimate only).
//   log10wOut (nPaths, nullable): per-path log10 weights.
int oracle_adjoint(const double *initialPos, uint32_t N, float arclength, uint32_t K, uint64_t seed,
      uint64_t nPaths, const float *table, int32_t tableSize, float ds, float minK, float maxK, float step,
      float absorption, int lerpMode, int threads, double *log10SumW, double *dTable, double *dSeg,
      double *log10wOut)
{
    if (N < 4 || K < 1 || tableSize < 2 || nPaths < 1) return 1;
    std::vector<V3<double>> init(N + 1);
    for (uint32_t i = 0; i <= N; ++i) init[i] = { initialPos[3 * i], initialPos[3 * i + 1], initialPos[3 * i + 2] };
    const AdjTable t { table, tableSize, minK, maxK, step, absorption, ds };
    const double invDs = (double)InvDs(N, arclength);
    const double absorbLn = (double)(N - 1) * std::log((double)std::exp((float)(-absorption * ds)));
    if (threads < 1) threads = 1;
    std::vector<double> lnw(nPaths);
    auto forRange = [&](auto fn) {
        std::vector<std::thread> pool;
        for (int th = 0; th < threads; ++th) {
            const uint64_t b = nPaths * (uint64_t)th / (uint64_t)threads;
            const uint64_t e = nPaths * (uint64_t)(th + 1) / (uint64_t)threads;
            pool.emplace_back([&, th, b, e] { fn(th, b, e); });
        }
        for (auto &p : pool) p.join();
    };
    forRange([&](int, uint64_t b, uint64_t e) {
        PathWork w;
        for (uint64_t p = b; p < e; ++p) lnw[p] = ForwardPath(init, N, invDs, K, seed, p, t, lerpMode, w);
    });
    long double M = -INFINITY;
    for (double v : lnw)
        if (!std::isnan(v)) M = std::max(M, (long double)v);
    long double S = 0.0L;
    for (double v : lnw)
        if (!std::isnan(v)) S += std::exp((long double)v - M);
    const double ln10 = std::log(10.0);
    if (log10SumW) *log10SumW = (double)((M + std::log(S) + absorbLn) / ln10);
    if (log10wOut)
        for (uint64_t p = 0; p < nPaths; ++p) log10wOut[p] = (lnw[p] + absorbLn) / ln10;
    if (!dTable && !dSeg) return 0;
    std::vector<std::vector<double>> dT(threads, std::vector<double>(tableSize, 0.0));
    std::vector<std::vector<double>> dS(threads, std::vector<double>(dSeg ? 3 * N : 0, 0.0));
    forRange([&](int th, uint64_t b, uint64_t e) {
        PathWork w;
        for (uint64_t p = b; p < e; ++p) {
            if (std::isnan(lnw[p])) continue;
            const double omega = (double)(std::exp((long double)lnw[p] - M) / S);
            if (!(omega > 0.0)) continue;
            ForwardPath(init, N, invDs, K, seed, p, t, lerpMode, w);
            AdjointPath(N, invDs, w, t, omega / ln10, dT[th].data(), dSeg ? dS[th].data() : nullptr);
        }
    });
    if (dTable) {
        for (int32_t b = 0; b < tableSize; ++b) {
            double s = 0.0;
            for (int th = 0; th < threads; ++th) s += dT[th][b];
            dTable[b] = s;
        }
    }
    if (dSeg) {
        for (uint32_t i = 0; i < 3 * N; ++i) {
            double s = 0.0;
            for (int th = 0; th < threads; ++th) s += dS[th][i];
            dSeg[i] = s;
        }
    }
    return 0;
}

}  // extern "C"
