// twisty_oracle.h -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Twisty's optimal-perturb path-integral hot path, used as the
// correctness checker for the HIP implementation and as the timed CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it (through oracle/_build/liboracle.so).  The shipped library never calls it.
//
// Every function cites the reference file:line it restates.  Reference paths are
// relative to the Twisty repository root.
//
// Pieces the reference does not define (it has no Philox, no usable initial-curve
// generator for these segment counts and no exit-direction histogram) are marked
// BUILD-DEFINED; they are specified here and pinned by closed forms in tests/.
//
// Geometry is templated on T (float = the reference's own precision for
// positions, double = the FP64 parity oracle).  The weight function keeps the
// reference's mixed precision exactly: float curvatures and float lerp, double
// log10 and double running sum (PathWeighters.cpp:27-104).
//
// Parity status per function (what pins it) is listed in the header of each
// section; a function without an independent pin says "parity unpinned".
#pragma once

#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace twisty_oracle {

constexpr float TwistyPi = 3.141592653589793f;  // MathConsts.h:8

// Pins the calling thread to the th-th CPU this process may run on (the CPU
// baseline runs one std::thread per allowed core, each on its own core).
inline void PinToAllowedCpu(int th)
{
    cpu_set_t allowed;
    CPU_ZERO(&allowed);
    if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return;
    const int n = CPU_COUNT(&allowed);
    if (n <= 0) return;
    int k = th % n;
    for (int c = 0; c < CPU_SETSIZE; ++c) {
        if (!CPU_ISSET(c, &allowed)) continue;
        if (k-- == 0) {
            cpu_set_t one;
            CPU_ZERO(&one);
            CPU_SET(c, &one);
            pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
            return;
        }
    }
}
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// A1. Counter-based draws (BUILD-DEFINED).
// The reference draws with std::mt19937_64 + libstdc++ distributions
// (FullExperimentRunnerOptimalPerturb.cpp:483-502); those are implementation
// defined, so both sides are fed Philox4x32-10 instead (Salmon, Moraes, Dror,
// Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11).
// Pinned by: published Philox4x32-10 known-answer vectors; exhaustive range
// enumeration of the index mapping at small N.
// ---------------------------------------------------------------------------
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

struct Draw {
    uint32_t diff;   // rightPointIndex - leftPointIndex
    uint32_t left;   // leftPointIndex
    float x;         // half rotation angle / pi, in (-1/2, 1/2); theta = 2*pi*x
};

// Draw for perturbation k of path `path` in stream `stream` (0 for a single
// experiment; the experiment index in batched runs).
// Counter = (k, stream, path_lo, path_hi), key = (seed_lo, seed_hi).
//   diff  ~ U{2 .. min(N-2, 25)}           FullExperimentRunnerOptimalPerturb.cpp:483-485
//   left  ~ U{1 .. N-diff-1}               FullExperimentRunnerOptimalPerturb.cpp:487-490
//   theta ~ U(-pi, pi)                     FullExperimentRunnerOptimalPerturb.cpp:501-502
// Integer ranges use the 32x32->64 multiply-shift map; x is an exactly
// representable float so CPU and GPU see identical angles.
inline Draw draw_perturbation(uint64_t seed, uint32_t stream, uint64_t path, uint32_t k, uint32_t N)
{
    const uint32_t ctr[4] = { k, stream, (uint32_t)path, (uint32_t)(path >> 32) };
    const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const uint32_t maxDiff = std::min<uint32_t>(N - 2, 25u);
    Draw d;
    d.diff = 2u + (uint32_t)(((uint64_t)r[0] * (uint64_t)(maxDiff - 1u)) >> 32);
    d.left = 1u + (uint32_t)(((uint64_t)r[1] * (uint64_t)(N - d.diff - 1u)) >> 32);
    const int32_t odd = (int32_t)(((r[2] >> 8) << 1) | 1u) - (int32_t)(1u << 24);
    d.x = (float)odd * 0x1p-25f;
    return d;
}

// ---------------------------------------------------------------------------
// Small vector type.
// ---------------------------------------------------------------------------
template<class T>
struct V3 {
    T x, y, z;
};
template<class T>
inline V3<T> sub(V3<T> a, V3<T> b)
{
    return { a.x - b.x, a.y - b.y, a.z - b.z };
}
template<class T>
inline V3<T> add(V3<T> a, V3<T> b)
{
    return { a.x + b.x, a.y + b.y, a.z + b.z };
}
template<class T>
inline V3<T> scale(V3<T> a, T s)
{
    return { a.x * s, a.y * s, a.z * s };
}
template<class T>
inline T dot(V3<T> a, V3<T> b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template<class T>
inline T norm(V3<T> a)
{
    return std::sqrt(dot(a, a));
}
// FMath's Vector3::Normalized is absent from the reference tree (fetched at
// configure time): v / |v| is assumed.  Parity unpinned below float rounding.
template<class T>
inline V3<T> normalized(V3<T> a)
{
    const T m = norm(a);
    return { a.x / m, a.y / m, a.z / m };
}

// ---------------------------------------------------------------------------
// A2/A3. Quaternion rotation.  CurvePerturbUtils.cpp:91-127.
// RotateQuatByQuat(a, b) returns the Hamilton product b (x) a; the sandwich
// below is therefore q (x) v (x) q*, an active rotation by +theta about n.
// Pinned by: Rodrigues' closed form, length preservation, theta and -theta.
// ---------------------------------------------------------------------------
template<class T>
inline void RotateQuatByQuat(const T *a, const T *b, T *out)
{
    const T t0 = (a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]);
    const T t1 = (a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2]);
    const T t2 = (a[0] * b[2] + a[1] * b[3] + a[2] * b[0] - a[3] * b[1]);
    const T t3 = (a[0] * b[3] - a[1] * b[2] + a[2] * b[1] + a[3] * b[0]);
    out[0] = t0;
    out[1] = t1;
    out[2] = t2;
    out[3] = t3;
}

template<class T>
inline void RotateVectorByQuaternion(const T *q, T *v)
{
    T p[4] = { T(0), v[0], v[1], v[2] };
    const T inv[4] = { q[0], -q[1], -q[2], -q[3] };
    RotateQuatByQuat(p, q, p);
    RotateQuatByQuat(inv, p, p);
    v[0] = p[1];
    v[1] = p[2];
    v[2] = p[3];
}

// One endpoint-preserving perturbation: FullExperimentRunnerOptimalPerturb.cpp:495-519.
// Points strictly between left and right rotate rigidly about the left point.
template<class T>
inline void ApplyPerturbation(V3<T> *pos, const Draw &d)
{
    const uint32_t left = d.left, right = d.left + d.diff;
    const V3<T> lp = pos[left];
    const V3<T> n = normalized(sub(pos[right], lp));
    const double halfAngle = (double)d.x;  // in units of pi
    const T s = (T)std::sin(kPi * halfAngle);
    const T c = (T)std::cos(kPi * halfAngle);
    const T q[4] = { c, n.x * s, n.y * s, n.z * s };
    for (uint32_t k = left + 1; k < right; ++k) {
        V3<T> sh = sub(pos[k], lp);
        T v[3] = { sh.x, sh.y, sh.z };
        RotateVectorByQuaternion(q, v);
        pos[k] = add(V3<T> { v[0], v[1], v[2] }, lp);
    }
}

// ---------------------------------------------------------------------------
// A4. Tangents.  CurvePerturbUtils.cpp:147-154 (host) / 197-216 (device form).
// A5. Joint curvature (radiative transfer).  CurvePerturbUtils.cpp:219-243.
// The host loop at CurvePerturbUtils.cpp:162 runs to numSegments and writes
// one element past the curvature array; the N-1 device loop is restated.
// invDs is computed in float exactly as the reference (uint32 / float).
// Pinned by: circular arc closed form 2 sin(alpha/2)/ds, straight line 0,
// reversal 2/ds.
// ---------------------------------------------------------------------------
template<class T>
inline void UpdateTangentsFromPos(const V3<T> *pos, V3<T> *tan, uint32_t N)
{
    for (uint32_t i = 0; i < N; ++i) tan[i] = normalized(sub(pos[i + 1], pos[i]));
}

inline float InvDs(uint32_t N, float arclength) { return (float)N / arclength; }

template<class T>
inline void UpdateCurvaturesFromTangents_RT(const V3<T> *tan, T *curv, uint32_t N, float arclength)
{
    const T invDs = (T)InvDs(N, arclength);
    for (uint32_t i = 0; i + 1 < N; ++i) {
        const V3<T> d = scale(sub(tan[i + 1], tan[i]), invDs);
        curv[i] = std::sqrt(dot(d, d));
    }
}

// ---------------------------------------------------------------------------
// A6. Per-path log10 weight.  PathWeighters.cpp:27-104, restated literally:
// snap within 1e-3 of the range ends (:45-51), clamp with a message (:53-65),
// floor index and last-bin fix (:67-74), lerp fraction in curvature units
// (:79-82, not divided by the step), per-segment absorption (:37,85), double
// log10 (:88) and NaN -> invalid (:89-102).  The message printing is replaced
// by a clamp counter.
// Pinned by: constant table -> (N-1) log10(c e^{-a ds}); linear table T[i]=i
// exposes the curvature-unit lerp exactly.
// ---------------------------------------------------------------------------
struct PathWeightValue {
    bool isValid = false;
    double weight = 0.0;
};

inline PathWeightValue WeightCurveViaCurvatureLog10(const float *curv, uint32_t numCurvatures,
      const float *table, int32_t tableSize, float ds, float minCurvature, float maxCurvature,
      float curvatureStepSize, float absorption, uint64_t *clampCount = nullptr)
{
    if (!curv || numCurvatures == 0) return { false, 0.0 };
    // std::exp of a float argument returns float (PathWeighters.cpp:37).
    const double absorptionFactor = (double)std::exp((float)(-absorption * ds));
    double running = 0.0;
    for (uint32_t seg = 0; seg < numCurvatures; ++seg) {
        float curvature = curv[seg];
        if ((curvature < minCurvature) && std::fabs(curvature - minCurvature) < 1e-3)
            curvature = minCurvature;
        if ((curvature > maxCurvature) && std::fabs(maxCurvature - curvature) < 1e-3)
            curvature = maxCurvature;
        if (curvature < minCurvature) {
            if (clampCount) ++*clampCount;
            curvature = minCurvature;
        }
        if (curvature > maxCurvature) {
            if (clampCount) ++*clampCount;
            curvature = maxCurvature;
        }
        const float distanceFromMin = curvature - minCurvature;
        const float realIdx = distanceFromMin / curvatureStepSize;
        int32_t leftIdx = (int32_t)std::floor(realIdx);
        int32_t rightIdx = leftIdx + 1;
        if (leftIdx == tableSize - 1) rightIdx--;
        // The reference reads table[tableSize] when realIdx rounds up to
        // tableSize; that cannot happen for curvatures <= maxCurvature with the
        // reference's own step, and is clamped here instead of read out of bounds.
        if (leftIdx > tableSize - 1) leftIdx = rightIdx = tableSize - 1;
        const float leftLookup = table[leftIdx];
        const float rightLookup = table[rightIdx];
        const float interpDist = distanceFromMin - ((float)leftIdx * curvatureStepSize);
        // float arithmetic, then promoted (PathWeighters.cpp:81-82); written as
        // separate statements so no FMA contraction changes the rounding.
        volatile float a = leftLookup * (1.0f - interpDist);
        volatile float b = rightLookup * interpDist;
        double interpolated = (double)(float)(a + b);
        interpolated *= absorptionFactor;
        const double l = std::log10(interpolated);
        if (std::isnan(l)) return { false, 0.0 };
        running += l;
    }
    if (std::isnan(running)) return { false, 0.0 };
    return { true, running };
}

// ---------------------------------------------------------------------------
// A7. Radiative-transfer weight table.  PathWeightUtils.cpp:19-33, 281-348,
// 350-433, restated literally, including
//   * IntegrandRT(p, kds, bds, eps, mu) being called with (mu, eps) swapped
//     (PathWeightUtils.cpp:401,404-405,412,420): the Gaussian phase width is the
//     config `eps`, the regulariser width is the config `mu`;
//   * the odd loop stopping at numStepsInt-3 (:409) and the even loop at
//     numStepsInt-4 (:417), so nodes numStepsInt-1 and numStepsInt-2 are zero;
//   * ascending sort before the sequential sum (:425-430);
//   * the float phase function and the float/double mix of every constant;
//   * entries <= 0 replaced by the smallest non-negative entry (:339-344).
// Pinned by: the scatter = 0 closed form
//   T(k) = bias/(2 pi^2) sqrt(pi/2) mu^-3 exp(-(k ds)^2 / (2 mu^2)),
// an independent FP64 Simpson of the same integrand, and kds -> 0 continuity.
// ---------------------------------------------------------------------------
struct WeightingParameters {  // PathWeightUtils.h:35-48
    float mu = 0.1f;
    uint32_t numStepsInt = 2000;
    float minBound = 0.0f;
    float maxBound = 100.0f;
    float eps = 0.01f;
    float bias = 1.0f;
    float scatter = 0.0f;
    float absorption = 0.0f;
    uint32_t numCurvatureSteps = 10000;
};

inline float SimpleGaussianPhase(float p, float mu)  // PathWeightUtils.cpp:19-23
{
    const float val = (-mu * p * p) * 0.5f;
    return std::exp(val);
}

inline float GaussianPhase(float evalLocation, float mu)  // PathWeightUtils.cpp:29-33
{
    const float Np = std::sqrt((TwistyPi * mu) * 0.5f) / (1.0f - std::exp(-2.0f / mu));
    return Np * SimpleGaussianPhase(evalLocation, mu);
}

// Argument names follow the declaration (PathWeightUtils.cpp:350).
inline double IntegrandRT(const double p, const double kds, const double bds, double eps, double mu)
{
    const double phaseFunction = GaussianPhase((float)p, (float)mu);
    const double scatteringTerm = p * std::exp(bds * phaseFunction - 1.0 * (eps * eps * p * p * 0.5));
    double sinTerm = p;
    if (kds > 0.0) sinTerm = std::sin(kds * p) / kds;
    return scatteringTerm * sinTerm;
}

// PathWeightUtils.cpp:371-433.  `scratch` avoids a per-call allocation.
inline float Integrate(float curvature, const WeightingParameters &wp, float ds,
      std::vector<double> &scratch)
{
    const double kds = (float)(curvature * ds);
    const double bds = (float)(wp.scatter * ds);
    const uint32_t n = wp.numStepsInt;
    const double h = (float)((wp.maxBound - wp.minBound) / (float)n);
    scratch.assign(n + 1, 0.0);
    const double weight = (double)(1.0f / 3.0f) * h;
    const double startPoint = wp.minBound;
    // (mu, eps) passed into the (eps, mu) slots, as the reference does.
    scratch[0] = IntegrandRT(startPoint, kds, bds, wp.mu, wp.eps) * weight;
    scratch[n] = IntegrandRT((double)wp.maxBound, kds, bds, wp.mu, wp.eps) * weight;
    for (uint32_t i = 1; i < n / 2; ++i) {
        const double evalP = startPoint + ((double)((2 * i) - 1) * h);
        scratch[(2 * i) - 1] = IntegrandRT(evalP, kds, bds, wp.mu, wp.eps) * weight * 4.0f;
    }
    for (uint32_t i = 1; i < (n / 2) - 1; ++i) {
        const double evalP = startPoint + ((double)(2 * i) * h);
        scratch[2 * i] = IntegrandRT(evalP, kds, bds, wp.mu, wp.eps) * weight * 2.0f;
    }
    std::sort(scratch.begin(), scratch.end());
    double firstVal = 0.0;
    for (const double v : scratch) firstVal += v;
    return (float)((firstVal / (2.0 * (double)TwistyPi * (double)TwistyPi)) * (double)wp.bias);
}

struct WeightTable {
    std::vector<float> table;
    float ds = 0.0f;
    float minCurvature = 0.0f;
    float maxCurvature = 0.0f;
    float curvatureStepSize = 0.0f;
    float minSegmentWeight = 0.0f;
    float maxSegmentWeight = 0.0f;
};

// Range and step: PathWeightUtils.cpp:37-53 and 281-289.
inline void TableRange(const WeightingParameters &wp, float ds, float &minK, float &maxK, float &step)
{
    minK = 0.0f;
    maxK = (float)((2.0 / (double)ds) * (double)1.01f);
    step = (maxK - minK) / (float)(wp.numCurvatureSteps - 1);
}

// Bins [begin, end) of WeightLookupTableIntegral::Initialize's loop
// (PathWeightUtils.cpp:316-332); callers may split the range across threads.
inline void IntegrateBins(const WeightingParameters &wp, float ds, uint32_t begin, uint32_t end,
      float *out)
{
    float minK, maxK, step;
    TableRange(wp, ds, minK, maxK, step);
    std::vector<double> scratch;
    for (uint32_t i = begin; i < end; ++i) {
        const float k = (i == 0) ? 0.0f : (minK + ((float)(int64_t)i * step));
        out[i] = Integrate(k, wp, ds, scratch);
    }
}

// Post-processing of Initialize (PathWeightUtils.cpp:303-347) given raw values.
inline void FinalizeTable(const std::vector<float> &raw, WeightTable &t)
{
    float mn = raw[0], mx = raw[0];
    for (size_t i = 1; i < raw.size(); ++i) {
        const float v = raw[i];
        if (v < mn && v >= 0.0f) mn = v;
        if (v > mx && v >= 0.0f) mx = v;
    }
    t.table = raw;
    for (float &v : t.table)
        if (v <= 0.0) v = mn;
    t.minSegmentWeight = mn;
    t.maxSegmentWeight = mx;
}

inline WeightTable BuildWeightTable(const WeightingParameters &wp, float ds)
{
    WeightTable t;
    t.ds = ds;
    TableRange(wp, ds, t.minCurvature, t.maxCurvature, t.curvatureStepSize);
    std::vector<float> raw(wp.numCurvatureSteps);
    IntegrateBins(wp, ds, 0, wp.numCurvatureSteps, raw.data());
    FinalizeTable(raw, t);
    return t;
}

// ---------------------------------------------------------------------------
// A8. Log-domain accumulator.  CombinedWeightUtils.h:22-33, 48-174.
// ExtractFinalValue's Boost cpp_dec_float_100 is replaced by returning the
// log10 of the value (long double), which keeps the same information without
// a big-float type; the last digits of Boost's rounding are parity unpinned.
// Pinned by: brute-force long double sum of 10^w, order independence,
// combine == sequential add.
// ---------------------------------------------------------------------------
constexpr double MaxDoubleLog10 = 300;
constexpr double MaxNumberOfPathsLog10 = 6;
constexpr uint32_t MaxNumPathsPerCombinedWeight = 1000000;

struct CombinedWeightValues_C {
    double m_maxWeightLog10 = 0.0;
    double m_runningTotal = 0.0;
    double m_offset = 0.0;
    uint32_t m_numValues = 0;
};

inline void CombinedWeightValues_C_Reset(CombinedWeightValues_C &c) { c = CombinedWeightValues_C {}; }

inline void CombinedWeightValues_C_AddValue(CombinedWeightValues_C &c, double valueLog10)
{
    if (c.m_numValues == 0) {
        c.m_maxWeightLog10 = valueLog10;
        c.m_offset = MaxDoubleLog10 - (c.m_maxWeightLog10 + MaxNumberOfPathsLog10);
        c.m_runningTotal += std::pow(10.0, (valueLog10 + c.m_offset));
        c.m_numValues++;
        return;
    }
    if (c.m_maxWeightLog10 >= valueLog10) {
        c.m_runningTotal += std::pow(10.0, (valueLog10 + c.m_offset));
        c.m_numValues++;
        return;
    }
    const double newOffset = MaxDoubleLog10 - (valueLog10 + MaxNumberOfPathsLog10);
    const double offsetDelta = newOffset - c.m_offset;
    const double log10RunningTotal = std::log10(c.m_runningTotal);
    c.m_runningTotal = std::pow(10.0, log10RunningTotal + offsetDelta);
    c.m_maxWeightLog10 = valueLog10;
    c.m_offset = newOffset;
    c.m_runningTotal += std::pow(10.0, (valueLog10 + c.m_offset));
    c.m_numValues++;
}

inline CombinedWeightValues_C CombinedWeightValues_C_CombineValues(
      CombinedWeightValues_C a, CombinedWeightValues_C b)
{
    if (a.m_numValues > 0 && b.m_numValues == 0) return a;
    if (b.m_numValues > 0 && a.m_numValues == 0) return b;
    if (b.m_maxWeightLog10 > a.m_maxWeightLog10) std::swap(a, b);
    CombinedWeightValues_C r;
    r.m_maxWeightLog10 = a.m_maxWeightLog10;
    r.m_numValues = a.m_numValues + b.m_numValues;
    r.m_offset = a.m_offset;
    if (a.m_maxWeightLog10 == b.m_maxWeightLog10) {
        r.m_runningTotal = a.m_runningTotal + b.m_runningTotal;
        return r;
    }
    const double offsetDelta = a.m_offset - b.m_offset;
    const double adj = std::pow(10.0, std::log10(b.m_runningTotal) + offsetDelta);
    r.m_runningTotal = a.m_runningTotal + adj;
    return r;
}

// log10 of ExtractFinalValue (CombinedWeightUtils.h:160-174); -inf when empty.
inline long double ExtractFinalValueLog10(const CombinedWeightValues_C &c)
{
    if (c.m_numValues == 0) return -std::numeric_limits<long double>::infinity();
    return (long double)std::log10(c.m_runningTotal) - (long double)c.m_offset;
}

inline long double LogSumExp10(long double a, long double b)
{
    if (std::isinf(a) && a < 0) return b;
    if (std::isinf(b) && b < 0) return a;
    const long double m = std::max(a, b);
    return m + std::log10(std::pow(10.0L, a - m) + std::pow(10.0L, b - m));
}

// ---------------------------------------------------------------------------
// A9. Path normaliser.  PathWeightUtils.cpp:643-752.  K4 closed form for N=4;
// N=5 and N=6 evaluate the nested 10000 x 1000 x 1000 quadrature (every K(N)
// macro calls F7, :703-705); everything else returns 1.0 (:745-750).
// The nested form is restated with its step counts as parameters so tests can
// run it small; the reference values are the defaults.
// Pinned by: the K4 closed form; for N = 5, 6, convergence at first order in
// the node spacing to (4 pi)^5 p_5(z) / ds^3, p_5 the closed-form 5-step
// random-flight density (each nesting level is a sphere convolution), with
// K5 == K6.  Boost big-float accumulation (long double here): parity unpinned.
// ---------------------------------------------------------------------------
inline long double K4(float z, float ds)  // PathWeightUtils.cpp:643-651
{
    float result = (z < 2.0) ? 1.0f : 0.0f;
    result /= z;
    result *= 2.0 * TwistyPi;
    const float invDs = (1.0 / ds);
    result *= (invDs * invDs * invDs);
    return result;
}

inline long double F5(float x, float z, float ds)  // :653-657
{
    const float rMag = std::sqrt((z * z) + 1.0 - (2.0 * z * x));
    return K4(rMag, ds);
}

template<int Order>
inline long double FN(float x, float z, float ds, uint32_t numSteps)  // F(N, M) macro :659-687
{
    const float rMag = std::sqrt((z * z) + 1.0 - (2.0 * z * x));
    long double result = 0.0;
    const float xMin = -1.0, xMax = 1.0;
    const float dx = (xMax - xMin) / numSteps;
    for (uint32_t i = 0; i < numSteps; i++) {
        const float xVal = xMin + i * dx;
        if constexpr (Order == 6)
            result += F5(xVal, rMag, ds);
        else
            result += FN<Order - 1>(xVal, rMag, ds, numSteps);
    }
    result *= (2.0 * TwistyPi) * dx;
    return result;
}

// K(N) macro (:689-711): every order integrates F7 over 10000 steps.
inline long double KMacro(double z, double ds, uint32_t outerSteps = 10000, uint32_t innerSteps = 1000)
{
    long double result = 0.0;
    const float xMin = -1.0, xMax = 1.0;
    const float dx = (xMax - xMin) / outerSteps;
    for (uint32_t i = 0; i < outerSteps; i++) {
        const float xVal = xMin + i * dx;
        result += FN<7>(xVal, (float)z, (float)ds, innerSteps);
    }
    result *= (2.0 * TwistyPi) * dx;
    return result;
}

struct BoundaryConditions {  // CurvePerturbUtils.h:62-68 / 70-76
    double startPos[3] = { 0, 0, 0 };
    double startDir[3] = { 1, 0, 0 };
    double endPos[3] = { 0, 0, 0 };
    double endDir[3] = { 1, 0, 0 };
    float arclength = 0.0f;
};

inline long double Norm(int numberOfSegments, float ds, const BoundaryConditions &bc,
      uint32_t outerSteps = 10000, uint32_t innerSteps = 1000)
{
    float zv[3];
    for (int a = 0; a < 3; ++a) {
        const float e = (float)bc.endPos[a], s = (float)bc.startPos[a];
        zv[a] = (e - s) * (1.0f / ds) - (float)bc.endDir[a] - (float)bc.startDir[a];
    }
    const float z = std::sqrt(zv[0] * zv[0] + zv[1] * zv[1] + zv[2] * zv[2]);
    switch (numberOfSegments) {
        case 4: return K4(z, ds);
        case 5:
        case 6: return KMacro(z, ds, outerSteps, innerSteps);
        default: return 1.0;
    }
}

// Coordinate axis least aligned with u (ties resolved x, then y, then z).
inline V3<double> LeastAlignedAxis(V3<double> u)
{
    const double ax = std::fabs(u.x), ay = std::fabs(u.y), az = std::fabs(u.z);
    int idx = 0;
    if (ay < ax) idx = 1;
    if (az < std::min(ax, ay)) idx = 2;
    if (idx == 0) return { 1, 0, 0 };
    if (idx == 1) return { 0, 1, 0 };
    return { 0, 0, 1 };
}

// ---------------------------------------------------------------------------
// A11. Initial curve (BUILD-DEFINED).
// Hinges follow Curve::SetBoundaryConditions (Curve.cpp:20-30):
//   p0 = start, p1 = start + ds*startDir, p_{N-1} = end - ds*endDir, p_N = end.
// The reference fills p1..p_{N-1} with ResolveEvenNumberOfSegments, which
// throws for these N (PathGeneration.cpp:236-247).  Here the M = N-2 free
// chords of length ds form a planar circular arc from p1 to p_{N-1}:
//   chord c = |p_{N-1} - p1| = 2R sin(M a/2), ds = 2R sin(a/2),
// solved for the turning angle a in [0, 2 pi/M) by bisection.  The arc bulges
// along startDir's component orthogonal to the chord (or the coordinate axis
// least aligned with the chord when that component vanishes).
// ds is the float value L/N, as the reference computes it (Curve.cpp:24).
// Pinned by: equal chord lengths, interior curvature 2 sin(a/2)/ds, endpoints,
// feasibility boundary = CalculateMinimumArclength (PathGeneration.cpp:12-43).
// Returns false when the arc does not exist (c > M ds or c == 0).
// ---------------------------------------------------------------------------
inline bool InitialArc(const BoundaryConditions &bc, uint32_t N, std::vector<V3<double>> &pos,
      double *turningAngle = nullptr, const V3<double> *bulge = nullptr)
{
    if (N < 3) return false;
    const double ds = (double)(bc.arclength / (float)N);
    const V3<double> s { bc.startPos[0], bc.startPos[1], bc.startPos[2] };
    const V3<double> sd { bc.startDir[0], bc.startDir[1], bc.startDir[2] };
    const V3<double> e { bc.endPos[0], bc.endPos[1], bc.endPos[2] };
    const V3<double> ed { bc.endDir[0], bc.endDir[1], bc.endDir[2] };
    pos.assign(N + 1, V3<double> { 0, 0, 0 });
    pos[0] = s;
    pos[1] = add(s, scale(sd, ds));
    pos[N - 1] = sub(e, scale(ed, ds));
    pos[N] = e;
    const uint32_t M = N - 2;
    const V3<double> chordVec = sub(pos[N - 1], pos[1]);
    const double c = norm(chordVec);
    if (!(c > 0.0)) return false;
    if (c > (double)M * ds * (1.0 + 1e-12)) return false;
    const V3<double> u = scale(chordVec, 1.0 / c);
    // bulge direction
    // (a given bulge direction replaces startDir here: the derivative in the end
    // direction is taken with the bulge held fixed when startDir lies on the chord)
    const V3<double> bsrc = bulge ? *bulge : sd;
    V3<double> v = sub(bsrc, scale(u, dot(bsrc, u)));
    if (norm(v) < 1e-9) {
        const V3<double> axis = LeastAlignedAxis(u);
        v = sub(axis, scale(u, dot(axis, u)));
    }
    v = scale(v, 1.0 / norm(v));
    const double target = c / ds;
    double alpha = 0.0;
    if (target < (double)M * (1.0 - 1e-15)) {
        // f(a) = sin(M a/2) / sin(a/2) decreases from M at a=0 to 0 at a=2 pi/M.
        double lo = 0.0, hi = 2.0 * kPi / (double)M;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            const double f = std::sin(0.5 * M * mid) / std::sin(0.5 * mid);
            if (f > target) lo = mid;
            else hi = mid;
        }
        alpha = 0.5 * (lo + hi);
    }
    if (turningAngle) *turningAngle = alpha;
    const V3<double> p1 = pos[1];
    if (alpha == 0.0) {
        for (uint32_t j = 1; j < M; ++j) pos[1 + j] = add(p1, scale(u, ds * (double)j));
    } else {
        const double R = ds / (2.0 * std::sin(0.5 * alpha));
        const double Phi = 0.5 * (double)M * alpha;
        const V3<double> mid = scale(add(pos[1], pos[N - 1]), 0.5);
        for (uint32_t j = 1; j < M; ++j) {
            const double phi = -Phi + (double)j * alpha;
            pos[1 + j] = add(mid, add(scale(u, R * std::sin(phi)), scale(v, R * (std::cos(phi) - std::cos(Phi)))));
        }
    }
    return true;
}

// Minimum arclength.  PathGeneration.cpp:12-43 (float arithmetic as written).
inline float CalculateMinimumArclength(const BoundaryConditions &bc, uint32_t numSegmentsPerCurve)
{
    const float M = (float)numSegmentsPerCurve;
    float S[3], D[3];
    for (int a = 0; a < 3; ++a) {
        S[a] = (float)bc.startDir[a] + (float)bc.endDir[a];
        D[a] = (float)bc.endPos[a] - (float)bc.startPos[a];
    }
    const float SS = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
    const float SD = S[0] * D[0] + S[1] * D[1] + S[2] * D[2];
    const float DD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    const float a = SS - ((M - 4.0f) * M + 4.0f);
    const float b = -2.0f * M * SD;
    const float c = M * M * DD;
    const float r1 = (-b - std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    const float r2 = (-b + std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    float m = 1000000.0f;
    if (!std::isnan(r1) && r1 > 0.0f) m = std::min(m, r1);
    if (!std::isnan(r2) && r2 > 0.0f) m = std::min(m, r2);
    return m;
}

// ---------------------------------------------------------------------------
// A10. Exit-direction histogram bin (BUILD-DEFINED; the reference has no such
// histogram -- its only histogram bins path-weight magnitudes,
// FullExperimentRunnerOptimalPerturb.cpp:26-37,562-575).
// The exit direction is the last free tangent t_{N-2}.  It is binned in
// (cos theta, phi) about the end direction e: cos theta = t.e in [-1, 1] over
// nTheta bins, phi = atan2(t.f2, t.f1) in [-pi, pi) over nPhi bins, with the
// end frame (f1, f2, e) built by EndFrame below.
// ---------------------------------------------------------------------------
inline void EndFrame(const double e[3], double f1[3], double f2[3])
{
    // f1 = normalize(a - (a.e) e) with a the coordinate axis least aligned with e.
    const V3<double> ax = LeastAlignedAxis(V3<double> { e[0], e[1], e[2] });
    const double a[3] = { ax.x, ax.y, ax.z };
    const double d = a[0] * e[0] + a[1] * e[1] + a[2] * e[2];
    double v[3] = { a[0] - d * e[0], a[1] - d * e[1], a[2] - d * e[2] };
    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) f1[i] = v[i] / n;
    f2[0] = e[1] * f1[2] - e[2] * f1[1];
    f2[1] = e[2] * f1[0] - e[0] * f1[2];
    f2[2] = e[0] * f1[1] - e[1] * f1[0];
}

inline uint32_t ExitBin(const double t[3], const double e[3], const double f1[3], const double f2[3],
      uint32_t nTheta, uint32_t nPhi)
{
    const double ct = t[0] * e[0] + t[1] * e[1] + t[2] * e[2];
    const double c1 = t[0] * f1[0] + t[1] * f1[1] + t[2] * f1[2];
    const double c2 = t[0] * f2[0] + t[1] * f2[1] + t[2] * f2[2];
    const double phi = std::atan2(c2, c1);
    int it = (int)std::floor((ct + 1.0) * 0.5 * nTheta);
    int ip = (int)std::floor((phi + kPi) / (2.0 * kPi) * nPhi);
    it = std::min(std::max(it, 0), (int)nTheta - 1);
    ip = std::min(std::max(ip, 0), (int)nPhi - 1);
    return (uint32_t)it * nPhi + (uint32_t)ip;
}

// ---------------------------------------------------------------------------
// One path end to end: K perturbations from the given state, then tangents,
// curvatures and weight (the per-path body of
// FullExperimentRunnerOptimalPerturb.cpp:469-593 with K perturbations before
// each weighting instead of one).
// ---------------------------------------------------------------------------
struct PathSpec {
    uint32_t N = 0;
    float arclength = 0.0f;
    uint32_t K = 1;
    uint64_t seed = 1;
    uint32_t stream = 0;
    uint32_t burnIn = 0;  // unweighted paths at the start of every chain
};

template<class T>
inline PathWeightValue WeightState(const std::vector<V3<T>> &pos, const PathSpec &ps,
      const WeightTable &wt, float absorption, std::vector<V3<T>> &tan, std::vector<T> &curvT,
      std::vector<float> &curvF, float *curvOut = nullptr, uint64_t *clampCount = nullptr)
{
    const uint32_t N = ps.N;
    tan.resize(N);
    curvT.resize(N - 1);
    curvF.resize(N - 1);
    UpdateTangentsFromPos(pos.data(), tan.data(), N);
    UpdateCurvaturesFromTangents_RT(tan.data(), curvT.data(), N, ps.arclength);
    for (uint32_t i = 0; i + 1 < N; ++i) curvF[i] = (float)curvT[i];
    if (curvOut)
        for (uint32_t i = 0; i + 1 < N; ++i) curvOut[i] = curvF[i];
    return WeightCurveViaCurvatureLog10(curvF.data(), N - 1, wt.table.data(), (int32_t)wt.table.size(),
          wt.ds, wt.minCurvature, wt.maxCurvature, wt.curvatureStepSize, absorption, clampCount);
}

}  // namespace twisty_oracle
