// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" wrappers over twisty_oracle.h for the Python tests (ctypes) and for
// bench.py's cpu_baseline leg.  Nothing in the shipped library links this.
#include "twisty_oracle.h"

#include <atomic>
#include <chrono>
#include <thread>

using namespace twisty_oracle;

namespace {

BoundaryConditions MakeBC(const double bc12[12], float arclength)
{
    BoundaryConditions bc;
    for (int a = 0; a < 3; ++a) {
        bc.startPos[a] = bc12[a];
        bc.startDir[a] = bc12[3 + a];
        bc.endPos[a] = bc12[6 + a];
        bc.endDir[a] = bc12[9 + a];
    }
    bc.arclength = arclength;
    return bc;
}

WeightingParameters MakeWP(float mu, float eps, float scatter, float absorption, float bias,
      uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound, float maxBound)
{
    WeightingParameters wp;
    wp.mu = mu;
    wp.eps = eps;
    wp.scatter = scatter;
    wp.absorption = absorption;
    wp.bias = bias;
    wp.numStepsInt = numStepsInt;
    wp.numCurvatureSteps = numCurvatureSteps;
    wp.minBound = minBound;
    wp.maxBound = maxBound;
    return wp;
}

struct SimTable {
    const float *table;
    int32_t size;
    float ds, minK, maxK, step, absorption;
};

// Per-thread accumulation: the reference's CombinedWeightValues_C per 1e6
// paths (FullExperimentRunnerOptimalPerturb.cpp:579-593) for sum w, plus a
// long double log-sum-exp of 2w for the Monte Carlo variance.
struct ThreadAcc {
    std::vector<CombinedWeightValues_C> acc;
    long double log10SumExact = -std::numeric_limits<long double>::infinity();
    long double log10SumW2 = -std::numeric_limits<long double>::infinity();
    long double maxW = -std::numeric_limits<long double>::infinity();
    uint64_t nValid = 0, nInvalid = 0, nClamped = 0;
    void add(double w)
    {
        if (acc.empty() || acc.back().m_numValues >= MaxNumPathsPerCombinedWeight) {
            acc.emplace_back();
            CombinedWeightValues_C_Reset(acc.back());
        }
        CombinedWeightValues_C_AddValue(acc.back(), w);
        log10SumExact = LogSumExp10(log10SumExact, (long double)w);
        log10SumW2 = LogSumExp10(log10SumW2, 2.0L * (long double)w);
        maxW = std::max(maxW, (long double)w);
        ++nValid;
    }
};

template<class T>
void SimulateRange(const std::vector<V3<double>> &init, const PathSpec &ps, uint32_t chainLength,
      uint64_t pathOffset, uint64_t first, uint64_t last, const SimTable &st, double *log10wOut,
      uint8_t *validOut, float *curvOut, float *exitOut, ThreadAcc &ta)
{
    const uint32_t N = ps.N;
    WeightTable wt;
    // Shallow view of the caller's table (copy is cheap next to the paths).
    wt.table.assign(st.table, st.table + st.size);
    wt.ds = st.ds;
    wt.minCurvature = st.minK;
    wt.maxCurvature = st.maxK;
    wt.curvatureStepSize = st.step;
    std::vector<V3<T>> initT(N + 1), pos(N + 1), tan;
    std::vector<T> curvT;
    std::vector<float> curvF;
    for (uint32_t i = 0; i <= N; ++i) initT[i] = { (T)init[i].x, (T)init[i].y, (T)init[i].z };
    for (uint64_t li = first; li < last; ++li) {
        const uint64_t path = pathOffset + li;
        if (li == first || (path % chainLength) == 0) {
            const uint64_t chainStart = path - (path % chainLength);
            pos = initT;
            // Burn-in (FullExperimentRunnerOptimalPerturb.cpp:469-478, 559-561):
            // ps.burnIn unweighted paths open every chain, drawn from path ids
            // 2^63 | (chain * burnIn + b).
            const uint64_t burnBase = (chainStart / chainLength) * (uint64_t)ps.burnIn;
            for (uint32_t b = 0; b < ps.burnIn; ++b)
                for (uint32_t k = 0; k < ps.K; ++k)
                    ApplyPerturbation(pos.data(),
                          draw_perturbation(ps.seed, ps.stream, (1ull << 63) | (burnBase + b), k, N));
            // Replay the chain from its start so any range split gives the same state.
            for (uint64_t p = chainStart; p < path; ++p)
                for (uint32_t k = 0; k < ps.K; ++k)
                    ApplyPerturbation(pos.data(), draw_perturbation(ps.seed, ps.stream, p, k, N));
        }
        for (uint32_t k = 0; k < ps.K; ++k)
            ApplyPerturbation(pos.data(), draw_perturbation(ps.seed, ps.stream, path, k, N));
        const PathWeightValue pw = WeightState(pos, ps, wt, st.absorption, tan, curvT, curvF,
              curvOut ? curvOut + li * (uint64_t)(N - 1) : nullptr, &ta.nClamped);
        if (log10wOut) log10wOut[li] = pw.weight;
        if (validOut) validOut[li] = pw.isValid ? 1 : 0;
        if (exitOut) {
            exitOut[li * 3 + 0] = (float)tan[N - 2].x;
            exitOut[li * 3 + 1] = (float)tan[N - 2].y;
            exitOut[li * 3 + 2] = (float)tan[N - 2].z;
        }
        if (pw.isValid) ta.add(pw.weight);
        else ++ta.nInvalid;
    }
}

}  // namespace

extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    philox4x32_10(ctr, key, out);
}

void oracle_draw(uint64_t seed, uint32_t stream, uint64_t path, uint32_t k, uint32_t N, uint32_t *diff,
      uint32_t *left, float *x)
{
    const Draw d = draw_perturbation(seed, stream, path, k, N);
    *diff = d.diff;
    *left = d.left;
    *x = d.x;
}

void oracle_rotate_vector_by_quaternion(const double q[4], double v[3]) { RotateVectorByQuaternion(q, v); }

void oracle_rotate_quat_by_quat(const double a[4], const double b[4], double out[4])
{
    RotateQuatByQuat(a, b, out);
}

// pos: (N+1)*3 doubles, perturbed in place.
void oracle_apply_perturbation_f64(double *pos, uint32_t left, uint32_t diff, float x)
{
    Draw d { diff, left, x };
    ApplyPerturbation(reinterpret_cast<V3<double> *>(pos), d);
}

// tangents (N*3) and curvatures (N-1) in FP64 from positions.
void oracle_tangents_curvatures_f64(const double *pos, uint32_t N, float arclength, double *tan,
      double *curv)
{
    UpdateTangentsFromPos(reinterpret_cast<const V3<double> *>(pos), reinterpret_cast<V3<double> *>(tan), N);
    UpdateCurvaturesFromTangents_RT(reinterpret_cast<const V3<double> *>(tan), curv, N, arclength);
}

float oracle_inv_ds(uint32_t N, float arclength) { return InvDs(N, arclength); }

int oracle_initial_arc(const double bc12[12], float arclength, uint32_t N, double *posOut, double *alphaOut)
{
    std::vector<V3<double>> pos;
    double alpha = 0.0;
    if (!InitialArc(MakeBC(bc12, arclength), N, pos, &alpha)) return 1;
    for (uint32_t i = 0; i <= N; ++i) {
        posOut[i * 3 + 0] = pos[i].x;
        posOut[i * 3 + 1] = pos[i].y;
        posOut[i * 3 + 2] = pos[i].z;
    }
    if (alphaOut) *alphaOut = alpha;
    return 0;
}

// InitialArc with the bulge direction taken from `bulge` instead of startDir.
int oracle_initial_arc_bulge(const double bc12[12], float arclength, uint32_t N, const double bulge[3],
      double *posOut)
{
    std::vector<V3<double>> pos;
    const V3<double> b { bulge[0], bulge[1], bulge[2] };
    if (!InitialArc(MakeBC(bc12, arclength), N, pos, nullptr, &b)) return 1;
    for (uint32_t i = 0; i <= N; ++i) {
        posOut[i * 3 + 0] = pos[i].x;
        posOut[i * 3 + 1] = pos[i].y;
        posOut[i * 3 + 2] = pos[i].z;
    }
    return 0;
}

float oracle_minimum_arclength(const double bc12[12], uint32_t N)
{
    return CalculateMinimumArclength(MakeBC(bc12, 0.0f), N);
}

double oracle_norm(uint32_t N, float ds, const double bc12[12], float arclength, uint32_t outerSteps,
      uint32_t innerSteps)
{
    return (double)Norm((int)N, ds, MakeBC(bc12, arclength), outerSteps, innerSteps);
}

float oracle_gaussian_phase(float p, float mu) { return GaussianPhase(p, mu); }

double oracle_integrand_rt(double p, double kds, double bds, double eps, double mu)
{
    return IntegrandRT(p, kds, bds, eps, mu);
}

void oracle_table_range(uint32_t numCurvatureSteps, float ds, float *minK, float *maxK, float *step)
{
    WeightingParameters wp;
    wp.numCurvatureSteps = numCurvatureSteps;
    TableRange(wp, ds, *minK, *maxK, *step);
}

// meta: [ds, minK, maxK, step, minSegmentWeight, maxSegmentWeight]
int oracle_build_table(float mu, float eps, float scatter, float absorption, float bias, uint32_t numStepsInt,
      uint32_t numCurvatureSteps, float minBound, float maxBound, float ds, int threads, float *tableOut,
      float *metaOut)
{
    if (numStepsInt < 4 || (numStepsInt % 2) != 0 || numCurvatureSteps < 2) return 1;
    const WeightingParameters wp = MakeWP(mu, eps, scatter, absorption, bias, numStepsInt,
          numCurvatureSteps, minBound, maxBound);
    WeightTable t;
    t.ds = ds;
    TableRange(wp, ds, t.minCurvature, t.maxCurvature, t.curvatureStepSize);
    std::vector<float> raw(numCurvatureSteps);
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    const uint32_t per = (numCurvatureSteps + threads - 1) / threads;
    for (int th = 0; th < threads; ++th) {
        const uint32_t b = std::min<uint32_t>(numCurvatureSteps, th * per);
        const uint32_t e = std::min<uint32_t>(numCurvatureSteps, b + per);
        pool.emplace_back([&, b, e] { IntegrateBins(wp, ds, b, e, raw.data()); });
    }
    for (auto &th : pool) th.join();
    FinalizeTable(raw, t);
    std::copy(t.table.begin(), t.table.end(), tableOut);
    const float meta[6] = { t.ds, t.minCurvature, t.maxCurvature, t.curvatureStepSize, t.minSegmentWeight,
        t.maxSegmentWeight };
    std::copy(meta, meta + 6, metaOut);
    return 0;
}

float oracle_integrate(float curvature, float mu, float eps, float scatter, float bias, uint32_t numStepsInt,
      float minBound, float maxBound, float ds)
{
    const WeightingParameters wp = MakeWP(mu, eps, scatter, 0.0f, bias, numStepsInt, 2, minBound, maxBound);
    std::vector<double> scratch;
    return Integrate(curvature, wp, ds, scratch);
}

// Returns 1 when valid; *w receives the log10 weight.
int oracle_weight_curvatures(const float *curv, uint32_t n, const float *table, int32_t size, float ds,
      float minK, float maxK, float step, float absorption, double *w)
{
    const PathWeightValue pw = WeightCurveViaCurvatureLog10(curv, n, table, size, ds, minK, maxK, step, absorption);
    *w = pw.weight;
    return pw.isValid ? 1 : 0;
}

// Accumulator over an array of log10 weights, 1e6 per accumulator as the
// reference.  out: [log10 total].
double oracle_accumulate_log10(const double *w, uint64_t n)
{
    std::vector<CombinedWeightValues_C> acc;
    for (uint64_t i = 0; i < n; ++i) {
        if (acc.empty() || acc.back().m_numValues >= MaxNumPathsPerCombinedWeight) acc.emplace_back();
        CombinedWeightValues_C_AddValue(acc.back(), w[i]);
    }
    long double tot = -std::numeric_limits<long double>::infinity();
    for (auto &a : acc) tot = LogSumExp10(tot, ExtractFinalValueLog10(a));
    return (double)tot;
}

// Pairwise combine of two accumulators built from w[0:m] and w[m:n].
double oracle_accumulate_combine_log10(const double *w, uint64_t m, uint64_t n)
{
    CombinedWeightValues_C a, b;
    for (uint64_t i = 0; i < m; ++i) CombinedWeightValues_C_AddValue(a, w[i]);
    for (uint64_t i = m; i < n; ++i) CombinedWeightValues_C_AddValue(b, w[i]);
    return (double)ExtractFinalValueLog10(CombinedWeightValues_C_CombineValues(a, b));
}

// Simulates paths [pathOffset, pathOffset + nPaths).
//   precision: 64 = FP64 oracle; 32 = reference-faithful float positions.
//   initialPos: (N+1)*3 doubles (the initial curve).
//   chainLength: 1 = every path restarts from initialPos; L > 1 = state is
//     carried through L consecutive path ids (the reference's per-thread chain,
//     FullExperimentRunnerOptimalPerturb.cpp:469-595).
//   burnIn: unweighted paths that open every chain (numPathsToSkip).
// Outputs (all nullable): log10w[nPaths], valid[nPaths], curv[nPaths*(N-1)],
//   exitDir[nPaths*3] (last free tangent).
// result: [log10 sum_valid 10^w through the reference accumulator,
//   log10 sum_valid 10^{2w}, max w, nValid, nInvalid, nClamped, seconds,
//   log10 sum_valid 10^w by long double log-sum-exp].  The two sums differ only
//   when an accumulator's first value is -inf (a zero-weight path): the
//   reference's AddValue then sets its offset to +inf and its total to NaN
//   (CombinedWeightUtils.h:62-69).
int oracle_simulate(int precision, const double *initialPos, uint32_t N, float arclength, uint32_t K,
      uint32_t chainLength, uint32_t burnIn, uint64_t seed, uint32_t stream, uint64_t pathOffset, uint64_t nPaths,
      const float *table, int32_t tableSize, float ds, float minK, float maxK, float step, float absorption,
      int threads, double *log10wOut, uint8_t *validOut, float *curvOut, float *exitOut, double *result)
{
    if (N < 4 || K < 1 || chainLength < 1 || tableSize < 2) return 1;
    std::vector<V3<double>> init(N + 1);
    for (uint32_t i = 0; i <= N; ++i) init[i] = { initialPos[3 * i], initialPos[3 * i + 1], initialPos[3 * i + 2] };
    PathSpec ps;
    ps.N = N;
    ps.arclength = arclength;
    ps.K = K;
    ps.seed = seed;
    ps.stream = stream;
    ps.burnIn = burnIn;
    const SimTable st { table, tableSize, ds, minK, maxK, step, absorption };
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > nPaths) threads = (int)std::max<uint64_t>(1, nPaths);
    std::vector<ThreadAcc> accs(threads);
    std::vector<std::thread> pool;
    const auto t0 = std::chrono::steady_clock::now();
    // Contiguous ranges keep the per-thread accumulator order deterministic.
    for (int th = 0; th < threads; ++th) {
        const uint64_t b = nPaths * (uint64_t)th / (uint64_t)threads;
        const uint64_t e = nPaths * (uint64_t)(th + 1) / (uint64_t)threads;
        pool.emplace_back([&, th, b, e] {
            PinToAllowedCpu(th);
            if (precision == 32)
                SimulateRange<float>(init, ps, chainLength, pathOffset, b, e, st, log10wOut, validOut, curvOut,
                      exitOut, accs[th]);
            else
                SimulateRange<double>(init, ps, chainLength, pathOffset, b, e, st, log10wOut, validOut, curvOut,
                      exitOut, accs[th]);
        });
    }
    for (auto &th : pool) th.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    long double tot = -std::numeric_limits<long double>::infinity();
    long double tot2 = -std::numeric_limits<long double>::infinity();
    long double totExact = -std::numeric_limits<long double>::infinity();
    long double mx = -std::numeric_limits<long double>::infinity();
    uint64_t nv = 0, ni = 0, nc = 0;
    for (auto &ta : accs) {
        for (auto &a : ta.acc) tot = LogSumExp10(tot, ExtractFinalValueLog10(a));
        tot2 = LogSumExp10(tot2, ta.log10SumW2);
        totExact = LogSumExp10(totExact, ta.log10SumExact);
        mx = std::max(mx, ta.maxW);
        nv += ta.nValid;
        ni += ta.nInvalid;
        nc += ta.nClamped;
    }
    if (result) {
        result[0] = (double)tot;
        result[1] = (double)tot2;
        result[2] = (double)mx;
        result[3] = (double)nv;
        result[4] = (double)ni;
        result[5] = (double)nc;
        result[6] = secs;
        result[7] = (double)totExact;
    }
    return 0;
}

void oracle_exit_bin(const double t[3], const double e[3], uint32_t nTheta, uint32_t nPhi, uint32_t *bin)
{
    double f1[3], f2[3];
    EndFrame(e, f1, f2);
    *bin = ExitBin(t, e, f1, f2, nTheta, nPhi);
}

void oracle_end_frame(const double e[3], double f1[3], double f2[3]) { EndFrame(e, f1, f2); }

}  // extern "C"
