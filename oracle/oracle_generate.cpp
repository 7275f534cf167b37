// oracle_generate.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 restatement of the independent-path estimator
// ExperimentBase::MSegmentPathGenerationMC (Experiments/ExperimentBase.cpp:
// 1246-1386): every path is built from scratch by recursive sphere
// intersection (ResolveEvenNumberOfSegments :806-983, ResolveThreeSegments
// :702-804, ResolveTwoSegments :617-698) between the hinge points, then weighted
// like the perturbed paths (tangents, curvatures, table lookup; :1318-1333).
//
// BUILD-DEFINED draws: the reference consumes one std::mt19937_64 per OpenMP
// thread in depth-first order, seeded from time(0) (:1277-1283), which no
// other implementation can reproduce.  Here every node of the recursion reads
// one Philox4x32-10 block keyed by the seed with counter
// (node, 0x80000000 | sub, path lo, path hi): node = 2^level + index in the
// level (heap numbering, leaves at the last level), sub = 1 for the two-segment
// step inside a three-segment leaf.  Words: u0 = phi, u1 = theta, bit 0 of the
// third word = the coin, u3 = the radius draw, each u = (word >> 8) * 2^-24.
// phi is sampled as cos(phi) = 1 - 2v, v = u0 * vmax, the reference's
// acos(1 - 2 U(0, vmax)) (:918-923, :778-780) written without the acos, with
// vmax = (1 - cos(phiExtent)) / 2 in closed form.
//
// Differences from the reference, on purpose: a recursion step that the
// reference aborts with an exception (spheres that do not meet, a NaN phi
// extent) marks the path invalid instead; the result keeps the reference's
// sum over valid paths and their count, min and max.
//
// The multi-arclength overload (:1388-1597) is oracle_generate_multi below, the
// two-table Metropolis-Hastings overload (:1599-1818, the NoisyCircle
// experiments' estimator) oracle_generate_chains.
//
// Pinned by (tests/test_oracle.py): exact segment lengths and hinge points of
// every generated path, the closed forms of the straight-line case, the
// exhaustive segment-count rule, the drawn theta and phi recovered from the
// geometry, (multi-ds) the table choice read back from flat tables, and
// (chains) the log against the natural log, the ds draw against the truncated
// exponential's distribution function, the sampling weight against its closed
// form, each proposal's table, weight and u read back from flat tables, and
// the chains against a replay of the acceptance rule.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

#include "twisty_oracle.h"

using namespace twisty_oracle;

namespace {

using V = V3<double>;

V cross(V a, V b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }

// Smallest distance of any branch decision of a path's recursion from its
// threshold (the 0.001 tolerances, the sphere-meeting tests, the frame's
// axis-choice limits).  Arithmetic of another precision can take the other
// branch only where this is small.
struct Margin {
    double m = INFINITY;
    void see(double q, double threshold) { m = std::min(m, std::fabs(q - threshold)); }
};
V mid(V a, V b) { return scale(add(a, b), 0.5); }
V nanv() { return { NAN, NAN, NAN }; }

struct GenWords {
    double u0, u1, u3;
    uint32_t coin;
};

GenWords gen_words(uint64_t seed, uint64_t path, uint32_t node, uint32_t sub)
{
    const uint32_t ctr[4] = { node, 0x80000000u | sub, (uint32_t)path, (uint32_t)(path >> 32) };
    const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const double s = 1.0 / 16777216.0;
    return { (double)(r[0] >> 8) * s, (double)(r[1] >> 8) * s, (double)(r[3] >> 8) * s, r[2] & 1u };
}

// Frame around z (ExperimentBase.cpp:868-875): x = z x e, y = z x x, with
// e = +x unless |z . x| > limit, then +y.
void frame(V z, double limit, V &x, V &y, Margin &mg)
{
    V other { 1.0, 0.0, 0.0 };
    mg.see(std::fabs(dot(z, other)), limit);
    if (std::fabs(dot(z, other)) > limit) other = { 0.0, 1.0, 0.0 };
    x = normalized(cross(z, other));
    y = normalized(cross(z, x));
}

V spherical(V x, V y, V z, double cosPhi, double sinPhi, double theta)
{
    return add(add(scale(x, sinPhi * std::cos(theta)), scale(y, sinPhi * std::sin(theta))), scale(z, cosPhi));
}

// phi with cos(phi) = 1 - 2 u vmax
void phi_of(double u, double vmax, double &c, double &s)
{
    const double v = u * vmax;
    c = 1.0 - 2.0 * v;
    s = 2.0 * std::sqrt(std::max(0.0, v * (1.0 - v)));
}

// ResolveTwoSegments (ExperimentBase.cpp:617-698): the middle point of two
// segments of length ds between L and R.
V resolve_two(V L, V R, double ds, const GenWords &g, Margin &mg)
{
    const double d = norm(sub(R, L));
    mg.see(std::fabs(2.0 * ds - d), 0.001);
    mg.see(d, 0.001);
    if (std::fabs(2.0 * ds - d) < 0.001) return mid(L, R);
    if (2.0 * ds < d) return nanv();
    const double theta = 2.0 * kPi * g.u1;
    if (d < 0.001) {
        double c, s;
        phi_of(g.u0, 1.0, c, s);
        return add(L, scale(V { s * std::cos(theta), s * std::sin(theta), c }, ds));
    }
    const V n = scale(sub(R, L), 1.0 / d);
    V other { 1.0, 0.0, 0.0 };
    mg.see(std::fabs(dot(n, other)), 0.99);
    if (std::fabs(dot(n, other)) >= 0.99) other = { 0.0, 1.0, 0.0 };
    const V m = normalized(cross(n, other));
    const double h = ds > 0.5 * d ? std::sqrt(ds * ds - 0.25 * d * d) : 0.0;
    // x_t = mid + h m rotated by theta about n (an active rotation, :688-697)
    return add(mid(L, R), scale(add(scale(m, std::cos(theta)), scale(cross(n, m), std::sin(theta))), h));
}

// ResolveThreeSegments (ExperimentBase.cpp:702-804): the first of the two
// points between L and R; the second follows from resolve_two.
V resolve_three_first(V L, V R, double ds, const GenWords &g, Margin &mg)
{
    const double d = norm(sub(R, L));
    mg.see(std::fabs(3.0 * ds - d), 0.001);
    mg.see(d, 0.001);
    if (std::fabs(3.0 * ds - d) < 0.001) return add(L, scale(sub(R, L), 1.0 / 3.0));
    if (3.0 * ds < d) return nanv();
    const double theta = 2.0 * kPi * g.u1;
    if (d < 0.001) {
        double c, s;
        phi_of(g.u0, 1.0, c, s);
        return add(L, scale(V { s * std::cos(theta), s * std::sin(theta), c }, ds));
    }
    const V z = scale(sub(R, L), 1.0 / d);
    V x, y;
    frame(z, 0.999, x, y, mg);
    const double r1 = ds, r2 = 2.0 * ds;
    double vmax;
    mg.see(d + r1, r2);
    if (d + r1 < r2) {
        vmax = 1.0;  // phiExtent = pi
    } else {
        const double h = 0.5 + (r1 * r1 - r2 * r2) / (2.0 * d * d);
        const double t = r1 * r1 - h * h * d * d;
        mg.see(std::fabs(t), 0.001);
        double a;
        if (std::fabs(t) < 0.001) a = 0.0;
        else if (t < 0.0) return nanv();  // the reference's NaN phi extent
        else a = std::sqrt(t);
        const double sE = std::min(1.0, a / r1);
        const double cE = std::sqrt(1.0 - sE * sE);
        // phiExtent = asin(sE), or pi - asin(sE) when h d < 0
        vmax = (h * d < 0.0) ? 0.5 + 0.5 * cE : 0.5 * sE * sE / (1.0 + cE);
    }
    double c, s;
    phi_of(g.u0, vmax, c, s);
    return add(L, scale(normalized(spherical(x, y, z, c, s, theta)), ds));
}

// ResolveEvenNumberOfSegments (ExperimentBase.cpp:806-983): the centre point
// of n segments between L and R.
V resolve_even_centre(uint32_t n, V L, V R, double ds, const GenWords &g, Margin &mg)
{
    const double d = norm(sub(R, L));
    mg.see(std::fabs((double)n * ds - d), 0.001);
    mg.see(d, 0.001);
    if (std::fabs((double)n * ds - d) < 0.001) return mid(L, R);
    const double rps = ds * (double)(n / 2);
    if (2.0 * rps < d) return nanv();
    const double theta = 2.0 * kPi * g.u1;
    V x, y;
    const V z = normalized(sub(R, L));
    frame(z, 0.999, x, y, mg);
    if (d < 0.001) {
        double c, s;
        phi_of(g.u0, 1.0, c, s);
        return add(L, scale(spherical(x, y, z, c, s, theta), rps * std::cbrt(g.u3)));
    }
    // a = sqrt(rps^2 - d^2/4), phiExtent = asin(a / rps): cos(phiExtent) = d / (2 rps)
    const double vmax = 0.5 * (1.0 - d / (2.0 * rps));
    double c, s;
    phi_of(g.u0, vmax, c, s);
    const double hyp = 0.5 * d / c;
    const double minPct = hyp / rps;
    const double radius = rps * std::cbrt(minPct + (1.0 - minPct) * g.u3);
    if (g.coin == 0) return add(L, scale(spherical(x, y, z, c, s, theta), radius));
    return add(R, scale(spherical(x, y, z, -c, s, theta), radius));
}

}  // namespace

// Levels of the recursion for N segments: N - 2 free segments halved until a
// span of 2 or 3 (each halving needs an even span); returns -1 when N is not of
// the form 2 + 2^k * {2, 3}, k >= 1 (the reference throws, :810-817).
int oracle_generate_levels(uint32_t N, uint32_t *leafSpan)
{
    if (N < 6) return -1;
    uint32_t span = N - 2;
    int k = 0;
    while (span > 3) {
        if (span & 1u) return -1;
        span >>= 1;
        ++k;
    }
    if (k < 1) return -1;
    *leafSpan = span;
    return k;
}

// One path's points (N+1) by the recursion; false when a step fails.
// margin (nullable): the smallest branch-decision margin of the path (Margin).
bool GeneratePath(const double bc12[12], double ds, uint32_t N, int levels, uint32_t leafSpan, uint64_t seed,
      uint64_t path, std::vector<V> &p, double *margin = nullptr)
{
    Margin mg;
    p.assign(N + 1, V { 0.0, 0.0, 0.0 });
    const V s0 { bc12[0], bc12[1], bc12[2] }, sd { bc12[3], bc12[4], bc12[5] };
    const V e0 { bc12[6], bc12[7], bc12[8] }, ed { bc12[9], bc12[10], bc12[11] };
    // hinge points (ExperimentBase.cpp:1255-1261)
    p[0] = s0;
    p[1] = add(s0, scale(sd, ds));
    p[N - 1] = sub(e0, scale(ed, ds));
    p[N] = e0;
    const uint32_t n0 = N - 2;
    for (int lv = 0; lv < levels; ++lv) {
        const uint32_t nodes = 1u << lv, span = n0 >> lv;
        for (uint32_t q = 0; q < nodes; ++q) {
            const uint32_t l = 1 + q * span;
            p[l + span / 2] =
                  resolve_even_centre(span, p[l], p[l + span], ds, gen_words(seed, path, nodes + q, 0), mg);
        }
    }
    const uint32_t nodes = 1u << levels;
    for (uint32_t q = 0; q < nodes; ++q) {
        const uint32_t l = 1 + q * leafSpan;
        if (leafSpan == 2) {
            p[l + 1] = resolve_two(p[l], p[l + 2], ds, gen_words(seed, path, nodes + q, 0), mg);
        } else {
            p[l + 1] = resolve_three_first(p[l], p[l + 3], ds, gen_words(seed, path, nodes + q, 0), mg);
            p[l + 2] = resolve_two(p[l + 1], p[l + 3], ds, gen_words(seed, path, nodes + q, 1), mg);
        }
    }
    if (margin) *margin = mg.m;
    for (const V &v : p)
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z)) return false;
    return true;
}

namespace {

// Build-defined natural logarithm of y in (0, 1] in FP64 from basic operations
// only, so that any IEEE double implementation forms the same bits (the
// chains' ds draw uses it): y = m 2^e with m in [sqrt(1/2), sqrt(2)), and
// ln m = 2 atanh(s) = 2 s (1 + z/3 + z^2/5 + ... + z^11/23), s = (m-1)/(m+1),
// z = s^2 <= 0.0295 (the next term is below 1e-18 of the sum).
double bd_log(double y)
{
    int e = 0;
    double m = std::frexp(y, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
    for (int k = 10; k >= 0; --k) p = std::fma(p, z, 1.0 / (double)(2 * k + 1));
    return std::fma((double)e, 0.69314718055994530942, (2.0 * s) * p);
}

// The chains' ds: minDs + x with x exponential of rate t truncated to
// [0, maxDs], by the inverse of its distribution function,
// x = -ln(1 - u (1 - exp(-t maxDs))) / t, u from node 0's first word; oneMinusQ
// = 1 - exp(-t maxDs) is formed once by the caller.  The reference draws
// exponentials and rejects those past maxDs + minDs (ExperimentBase.cpp:
// 1662-1667): the same distribution.
float chain_ds(double u, double oneMinusQ, float t, float minDs)
{
    const double y = std::fma(-u, oneMinusQ, 1.0);
    const double x = -bd_log(y) / (double)t;
    return (float)((double)minDs + x);
}

// log10 of the reference's sampling weight 1 / (t exp(-t (ds - minDs) N)),
// formed in float as ExperimentBase.cpp:1669-1670 forms it.
double chain_log10_sampling_weight(float t, float ds, float minDs, uint32_t N)
{
    const float sw = 1.0f / (t * std::exp(-t * (ds - minDs) * (float)N));
    return std::log10((double)sw);
}

// The table of a multi-ds set whose ds is nearest (first on ties,
// PathWeightUtils.cpp:223-236).
uint32_t nearest_table(const std::vector<WeightTable> &wts, float ds)
{
    uint32_t best = 0;
    float bestDist = std::numeric_limits<float>::max();
    for (uint32_t t = 0; t < (uint32_t)wts.size(); ++t) {
        const float dist = std::fabs(wts[t].ds - ds);
        if (dist < bestDist) {
            bestDist = dist;
            best = t;
        }
    }
    return best;
}

std::vector<WeightTable> multi_tables(const float *tables, int32_t bins, const float *meta, uint32_t numTables)
{
    std::vector<WeightTable> wts(numTables);
    for (uint32_t t = 0; t < numTables; ++t) {
        wts[t].table.assign(tables + (uint64_t)t * bins, tables + (uint64_t)(t + 1) * bins);
        wts[t].ds = meta[4 * t];
        wts[t].minCurvature = meta[4 * t + 1];
        wts[t].maxCurvature = meta[4 * t + 2];
        wts[t].curvatureStepSize = meta[4 * t + 3];
    }
    return wts;
}

}  // namespace

extern "C" {

int oracle_generate_levels_c(uint32_t N, uint32_t *leafSpan) { return oracle_generate_levels(N, leafSpan); }

// Paths [pathOffset, pathOffset + nPaths) of the independent-path estimator.
//   bc12: start, start dir, end, end dir (unit); arclength: the boundary's; ds =
//   arclength / N in float (ExperimentBase.cpp:1253).
// Outputs (nullable): log10w[n] (path log10 weight without normaliser; NaN when
// invalid), valid[n], positions[n*(N+1)*3], curv[n*(N-1)] (the float joint
// curvatures the lookup uses; written for valid paths), margin[n] (the path's
// smallest branch-decision margin, Margin).
// result: [log10 sum_valid 10^w (long double LSE), nValid, nInvalid, min w,
//   max w, seconds].
int oracle_generate(const double bc12[12], float arclength, uint32_t N, uint64_t seed, uint64_t pathOffset,
      uint64_t nPaths, const float *table, int32_t tableSize, float ds, float minK, float maxK, float step,
      float absorption, int threads, double *log10wOut, uint8_t *validOut, double *positionsOut, float *curvOut,
      double *marginOut, double *result)
{
    uint32_t leafSpan = 0;
    const int levels = oracle_generate_levels(N, &leafSpan);
    if (levels < 0 || tableSize < 2 || !(arclength > 0.0f)) return 1;
    const double dsGen = (double)(arclength / (float)N);
    WeightTable wt;
    wt.table.assign(table, table + tableSize);
    wt.ds = ds;
    wt.minCurvature = minK;
    wt.maxCurvature = maxK;
    wt.curvatureStepSize = step;
    PathSpec ps;
    ps.N = N;
    ps.arclength = arclength;
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nPaths) threads = (int)std::max<uint64_t>(1, nPaths);
    struct Acc {
        long double sum = -std::numeric_limits<long double>::infinity();
        uint64_t nv = 0, ni = 0;
        double mn = INFINITY, mx = -INFINITY;
    };
    std::vector<Acc> accs(threads);
    std::vector<std::thread> pool;
    const auto t0 = std::chrono::steady_clock::now();
    for (int th = 0; th < threads; ++th) {
        const uint64_t b = nPaths * (uint64_t)th / (uint64_t)threads;
        const uint64_t e = nPaths * (uint64_t)(th + 1) / (uint64_t)threads;
        pool.emplace_back([&, th, b, e] {
            PinToAllowedCpu(th);
            std::vector<V> p;
            std::vector<V> tan;
            std::vector<double> curvT;
            std::vector<float> curvF;
            Acc &a = accs[th];
            for (uint64_t li = b; li < e; ++li) {
                const bool ok = GeneratePath(bc12, dsGen, N, levels, leafSpan, seed, pathOffset + li, p,
                      marginOut ? marginOut + li : nullptr);
                if (positionsOut)
                    for (uint32_t i = 0; i <= N; ++i) {
                        positionsOut[(li * (N + 1) + i) * 3 + 0] = p[i].x;
                        positionsOut[(li * (N + 1) + i) * 3 + 1] = p[i].y;
                        positionsOut[(li * (N + 1) + i) * 3 + 2] = p[i].z;
                    }
                PathWeightValue pw { false, NAN };
                if (ok)
                    pw = WeightState(p, ps, wt, absorption, tan, curvT, curvF,
                          curvOut ? curvOut + li * (uint64_t)(N - 1) : nullptr);
                if (log10wOut) log10wOut[li] = pw.isValid ? pw.weight : NAN;
                if (validOut) validOut[li] = pw.isValid ? 1 : 0;
                if (!pw.isValid) {
                    ++a.ni;
                    continue;
                }
                ++a.nv;
                a.sum = LogSumExp10(a.sum, (long double)pw.weight);
                a.mn = std::min(a.mn, pw.weight);
                a.mx = std::max(a.mx, pw.weight);
            }
        });
    }
    for (auto &t : pool) t.join();
    Acc tot;
    for (const Acc &a : accs) {
        tot.sum = LogSumExp10(tot.sum, a.sum);
        tot.nv += a.nv;
        tot.ni += a.ni;
        tot.mn = std::min(tot.mn, a.mn);
        tot.mx = std::max(tot.mx, a.mx);
    }
    result[0] = (double)tot.sum;
    result[1] = (double)tot.nv;
    result[2] = (double)tot.ni;
    result[3] = tot.mn;
    result[4] = tot.mx;
    result[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// The multi-arclength form (ExperimentBase.cpp:1388-1597): per path ds =
// u (maxDs - minDs) + minDs in float (u from node 0's words), the table whose
// ds is nearest (first on ties, PathWeightUtils.cpp:223-236), curvature with
// arclength ds N, absorption with the table's ds.  tables: T x bins floats;
// meta: 4 per table (ds, minK, maxK, step).  Outputs as oracle_generate, plus
// tableOut[n] (nullable): the table each path is weighted with.
int oracle_generate_multi(const double bc12[12], uint32_t N, uint64_t seed, uint64_t pathOffset, uint64_t nPaths,
      const float *tables, int32_t bins, const float *meta, uint32_t numTables, float absorption, float minDs,
      float maxDs, int threads, double *log10wOut, uint8_t *validOut, float *curvOut, double *marginOut,
      uint32_t *tableOut, double *result)
{
    uint32_t leafSpan = 0;
    const int levels = oracle_generate_levels(N, &leafSpan);
    if (levels < 0 || bins < 2 || numTables < 1 || !(minDs > 0.0f) || !(maxDs >= minDs)) return 1;
    const std::vector<WeightTable> wts = multi_tables(tables, bins, meta, numTables);
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nPaths) threads = (int)std::max<uint64_t>(1, nPaths);
    struct Acc {
        long double sum = -std::numeric_limits<long double>::infinity();
        uint64_t nv = 0, ni = 0;
        double mn = INFINITY, mx = -INFINITY;
    };
    std::vector<Acc> accs(threads);
    std::vector<std::thread> pool;
    const auto t0 = std::chrono::steady_clock::now();
    for (int th = 0; th < threads; ++th) {
        const uint64_t b = nPaths * (uint64_t)th / (uint64_t)threads;
        const uint64_t e = nPaths * (uint64_t)(th + 1) / (uint64_t)threads;
        pool.emplace_back([&, th, b, e] {
            PinToAllowedCpu(th);
            std::vector<V> p;
            std::vector<V> tan;
            std::vector<double> curvT;
            std::vector<float> curvF;
            Acc &a = accs[th];
            for (uint64_t li = b; li < e; ++li) {
                const uint64_t path = pathOffset + li;
                const float u = (float)gen_words(seed, path, 0, 0).u0;
                const float ds = std::fma(u, maxDs - minDs, minDs);
                const uint32_t best = nearest_table(wts, ds);
                PathSpec ps;
                ps.N = N;
                ps.arclength = ds * (float)N;
                if (tableOut) tableOut[li] = best;
                const bool ok = GeneratePath(bc12, (double)ds, N, levels, leafSpan, seed, path, p,
                      marginOut ? marginOut + li : nullptr);
                PathWeightValue pw { false, NAN };
                if (ok)
                    pw = WeightState(p, ps, wts[best], absorption, tan, curvT, curvF,
                          curvOut ? curvOut + li * (uint64_t)(N - 1) : nullptr);
                if (log10wOut) log10wOut[li] = pw.isValid ? pw.weight : NAN;
                if (validOut) validOut[li] = pw.isValid ? 1 : 0;
                if (!pw.isValid) {
                    ++a.ni;
                    continue;
                }
                ++a.nv;
                a.sum = LogSumExp10(a.sum, (long double)pw.weight);
                a.mn = std::min(a.mn, pw.weight);
                a.mx = std::max(a.mx, pw.weight);
            }
        });
    }
    for (auto &t : pool) t.join();
    Acc tot;
    for (const Acc &a : accs) {
        tot.sum = LogSumExp10(tot.sum, a.sum);
        tot.nv += a.nv;
        tot.ni += a.ni;
        tot.mn = std::min(tot.mn, a.mn);
        tot.mx = std::max(tot.mx, a.mx);
    }
    result[0] = (double)tot.sum;
    result[1] = (double)tot.nv;
    result[2] = (double)tot.ni;
    result[3] = tot.mn;
    result[4] = tot.mx;
    result[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// The two-table Metropolis-Hastings overload (ExperimentBase.cpp:1599-1818),
// the estimator of the NoisyCircle experiments, with chains in place of the
// reference's OpenMP threads: chains [chainOffset, chainOffset + nChains), each
// of chainLength proposals; proposal s of chain c draws as path
// c * chainLength + s.  Per proposal (:1658-1720): ds = chain_ds (node 0's first
// word, minDs and maxDs as the reference's minDs and actualMaxDs), the table
// nearest to it, the path built with that ds and weighted as the perturbed
// paths are, with absorption + scatter as the absorption coefficient
// (WeightCurveViaPositionLog10_PositionDependent, PathWeighters.cpp:105-177: its
// object-table branch is commented out, so every joint takes the environment
// table and the loss exp(-(absorption + scatter) ds) of the table's ds), plus
// log10 of the sampling weight; extinction = absorption + scatter in float, the
// rate of the ds draw too.  An invalid proposal is skipped (:1720-1723); a
// chain's first valid proposal is accepted, a later one when u <= 10^w / 10^w'
// with the normaliser in both and the ratio in FP64 as pow(10, w)
// (:1730-1755; u from node 0's second word), and every valid proposal adds the
// chain's current weight to the sum (:1757-1784).
// Outputs per proposal (nullable): log10wOut (log10 weight with the sampling
// weight, without the normaliser; NaN when invalid), validOut, chainOut (the
// chain's current log10 weight after the proposal, without the normaliser; NaN
// before its first valid one), curvOut, marginOut (branch margin of the path),
// acceptMarginOut (|u - ratio| / ratio of the acceptance test; +inf where there
// was none), uOut (u), dsOut (ds), tableOut.  result: [log10 of the sum of 10^(chain weight +
// normaliser) over valid proposals, nValid, nInvalid, min and max of those
// chain weights with the normaliser, seconds].
int oracle_generate_chains(const double bc12[12], uint32_t N, uint64_t seed, uint64_t chainOffset, uint64_t nChains,
      uint64_t chainLength, const float *tables, int32_t bins, const float *meta, uint32_t numTables,
      float extinction, float minDs, float maxDs, double normLog10, int threads, double *log10wOut,
      uint8_t *validOut, double *chainOut, float *curvOut, double *marginOut, double *acceptMarginOut,
      double *uOut, float *dsOut, uint32_t *tableOut, double *result)
{
    uint32_t leafSpan = 0;
    const int levels = oracle_generate_levels(N, &leafSpan);
    if (levels < 0 || bins < 2 || numTables < 1 || !(minDs > 0.0f) || !(maxDs > 0.0f) || !(extinction > 0.0f) ||
          chainLength < 1 || !((double)extinction * (double)maxDs < 700.0))
        return 1;
    const std::vector<WeightTable> wts = multi_tables(tables, bins, meta, numTables);
    const double oneMinusQ = 1.0 - std::exp(-(double)extinction * (double)maxDs);
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nChains) threads = (int)std::max<uint64_t>(1, nChains);
    struct Acc {
        long double sum = -std::numeric_limits<long double>::infinity();
        uint64_t nv = 0, ni = 0;
        double mn = INFINITY, mx = -INFINITY;
    };
    std::vector<Acc> accs(threads);
    std::vector<std::thread> pool;
    const auto t0 = std::chrono::steady_clock::now();
    for (int th = 0; th < threads; ++th) {
        const uint64_t b = nChains * (uint64_t)th / (uint64_t)threads;
        const uint64_t e = nChains * (uint64_t)(th + 1) / (uint64_t)threads;
        pool.emplace_back([&, th, b, e] {
            PinToAllowedCpu(th);
            std::vector<V> p;
            std::vector<V> tan;
            std::vector<double> curvT;
            std::vector<float> curvF;
            Acc &a = accs[th];
            for (uint64_t lc = b; lc < e; ++lc) {
                bool first = true;
                double prev = NAN;  // the chain's current weight with the normaliser
                for (uint64_t st = 0; st < chainLength; ++st) {
                    const uint64_t li = lc * chainLength + st;
                    const uint64_t path = (chainOffset + lc) * chainLength + st;
                    const GenWords w0 = gen_words(seed, path, 0, 0);
                    const float ds = chain_ds(w0.u0, oneMinusQ, extinction, minDs);
                    const uint32_t best = nearest_table(wts, ds);
                    PathSpec ps;
                    ps.N = N;
                    ps.arclength = ds * (float)N;
                    if (tableOut) tableOut[li] = best;
                    if (uOut) uOut[li] = w0.u1;
                    if (dsOut) dsOut[li] = ds;
                    if (acceptMarginOut) acceptMarginOut[li] = INFINITY;
                    const bool ok = GeneratePath(bc12, (double)ds, N, levels, leafSpan, seed, path, p,
                          marginOut ? marginOut + li : nullptr);
                    PathWeightValue pw { false, NAN };
                    if (ok)
                        pw = WeightState(p, ps, wts[best], extinction, tan, curvT, curvF,
                              curvOut ? curvOut + li * (uint64_t)(N - 1) : nullptr);
                    if (pw.isValid) pw.weight += chain_log10_sampling_weight(extinction, ds, minDs, N);
                    if (log10wOut) log10wOut[li] = pw.isValid ? pw.weight : NAN;
                    if (validOut) validOut[li] = pw.isValid ? 1 : 0;
                    if (!pw.isValid) {
                        ++a.ni;
                        if (chainOut) chainOut[li] = first ? NAN : prev - normLog10;
                        continue;
                    }
                    const double cur = pw.weight + normLog10;
                    if (first) {
                        first = false;
                        prev = cur;
                    } else {
                        const double ratio = std::pow(10.0, cur) / std::pow(10.0, prev);
                        const float u = (float)w0.u1;
                        if (acceptMarginOut) acceptMarginOut[li] = std::fabs((double)u - ratio) / ratio;
                        if ((double)u <= ratio) prev = cur;
                    }
                    if (chainOut) chainOut[li] = prev - normLog10;
                    ++a.nv;
                    a.sum = LogSumExp10(a.sum, (long double)prev);
                    a.mn = std::min(a.mn, prev);
                    a.mx = std::max(a.mx, prev);
                }
            }
        });
    }
    for (auto &t : pool) t.join();
    Acc tot;
    for (const Acc &a : accs) {
        tot.sum = LogSumExp10(tot.sum, a.sum);
        tot.nv += a.nv;
        tot.ni += a.ni;
        tot.mn = std::min(tot.mn, a.mn);
        tot.mx = std::max(tot.mx, a.mx);
    }
    result[0] = (double)tot.sum;
    result[1] = (double)tot.nv;
    result[2] = (double)tot.ni;
    result[3] = tot.mn;
    result[4] = tot.mx;
    result[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// The build-defined pieces of the chains' draws, for the pins.
double oracle_bd_log(double y) { return bd_log(y); }
float oracle_chain_ds(double u, float extinction, float minDs, float maxDs)
{
    return chain_ds(u, 1.0 - std::exp(-(double)extinction * (double)maxDs), extinction, minDs);
}
double oracle_chain_log10_sampling_weight(float extinction, float ds, float minDs, uint32_t N)
{
    return chain_log10_sampling_weight(extinction, ds, minDs, N);
}

}  // extern "C"
