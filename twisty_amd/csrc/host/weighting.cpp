// weighting.cpp -- curvature -> segment-weight tables (PathWeightUtils.cpp).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <memory>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "twisty_host.hpp"

namespace twisty {

namespace {
constexpr float TwistyPi = 3.141592653589793f;  // MathConsts.h:8

uint64_t fnv1a(const std::string &s)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string hex_of(float v)
{
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    std::ostringstream os;
    os << "0x" << std::uppercase << std::setfill('0') << std::setw(8) << std::hex << bits;
    return os.str();
}

int hardware_threads()
{
    const unsigned n = std::thread::hardware_concurrency();
    return n ? (int)n : 1;
}
}  // namespace

std::pair<std::string, uint64_t> WeightingParameters::GenerateStringUUID() const
{
    std::ostringstream u;
    u << "mu_" << hex_of(mu) << "_numStepsInt_" << numStepsInt << "_minBound_" << hex_of(minBound) << "_maxBound_"
      << hex_of(maxBound) << "_eps_" << hex_of(eps) << "_scatter_" << hex_of(scatter) << "_absorption_"
      << hex_of(absorption) << "_bias_" << hex_of(bias) << "_numCurvatureSteps_" << numCurvatureSteps
      << "_weightingMethod_" << (int32_t)weightingMethod;
    return { u.str(), fnv1a(u.str()) };
}

namespace PathWeighting {

// PathWeightUtils.cpp:19-33, float arithmetic as written.
float GaussianPhase(float evalLocation, float mu)
{
    const float Np = std::sqrt((TwistyPi * mu) * 0.5f) / (1.0f - std::exp(-2.0f / mu));
    const float val = (-mu * evalLocation * evalLocation) * 0.5f;
    return Np * std::exp(val);
}

// PathWeightUtils.cpp:350-369.  Parameter names as declared; callers pass the
// config's (mu, eps) into (eps, mu), as the reference's Integrate does.
double IntegrandRT(double p, double kds, double bds, double eps, double mu)
{
    const double phaseFunction = GaussianPhase((float)p, (float)mu);
    const double scatteringTerm = p * std::exp(bds * phaseFunction - 1.0 * (eps * eps * p * p * 0.5));
    const double sinTerm = (kds > 0.0) ? std::sin(kds * p) / kds : p;
    return scatteringTerm * sinTerm;
}

BaseWeightLookupTable::BaseWeightLookupTable(const WeightingParameters &wp, float ds, float minCurvature,
      float maxCurvature)
    : m_weightingParams(wp)
    , m_ds(ds)
    , m_minCurvature(minCurvature)
    , m_maxCurvature(maxCurvature)
    , m_curvatureStepSize((maxCurvature - minCurvature) / (float)(wp.numCurvatureSteps - 1))
    , m_lookupTable(wp.numCurvatureSteps)
{
}

void BaseWeightLookupTable::StreamOut(std::ostream &os) const
{
    os.write(reinterpret_cast<const char *>(m_lookupTable.data()), sizeof(float) * m_lookupTable.size());
    os.write(reinterpret_cast<const char *>(&m_minSegmentWeight), sizeof(float));
    os.write(reinterpret_cast<const char *>(&m_maxSegmentWeight), sizeof(float));
}

void BaseWeightLookupTable::InitializeFromStream(std::istream &is)
{
    m_lookupTable.resize(m_weightingParams.numCurvatureSteps);
    is.read(reinterpret_cast<char *>(m_lookupTable.data()), sizeof(float) * m_lookupTable.size());
    is.read(reinterpret_cast<char *>(&m_minSegmentWeight), sizeof(float));
    is.read(reinterpret_cast<char *>(&m_maxSegmentWeight), sizeof(float));
    if (!is) throw std::runtime_error("weight table stream ended early");
}

std::string BaseWeightLookupTable::UUID() const
{
    return "ds_" + hex_of(m_ds) + "wp_" + m_weightingParams.GenerateStringUUID().first;
}

void BaseWeightLookupTable::ExportValues(const std::string &directory) const
{
    const std::string name = directory + "/ds_" + hex_of(m_ds) + "_" + std::to_string(fnv1a(UUID())) + "_" +
          ExportFilename();
    std::ofstream out(name);
    if (!out.is_open()) throw std::runtime_error("Failed to export weight table: " + name);
    for (size_t i = 0; i < m_lookupTable.size(); ++i)
        out << (double)(m_minCurvature + (float)i * m_curvatureStepSize) << ", " << m_lookupTable[i] << '\n';
}

// Range: PathWeightUtils.cpp:281-289 (maxK = (2.0 / ds) * 1.01f in double,
// stored as float).
WeightLookupTableIntegral::WeightLookupTableIntegral(const WeightingParameters &wp, float ds)
    : BaseWeightLookupTable(wp, ds, 0.0f, (float)((2.0 / (double)ds) * (double)1.01f))
{
}

// Simpson rule of PathWeightUtils.cpp:371-433 with its exact node set and
// rounding: h in float, weights (1/3)h, 4(1/3)h, 2(1/3)h as
// (IntegrandRT * weight) * 4.0f, odd nodes 1..n-3 and even nodes 2..n-4 (the
// loops stop early, so nodes n-1 and n-2 contribute nothing), node n at
// maxBound, ascending sort, sequential double sum, result / (2 pi^2) * bias in
// float.  The factor p*exp(bds*phase - c^2 p^2/2) of every node is computed once
// per table (it does not depend on the curvature); a node whose factor is
// exactly zero contributes exactly +-0 and is left out of the sorted sum,
// which does not change any partial sum.
namespace {
struct NodeSet {
    std::vector<double> p;       // node abscissae
    std::vector<double> factor;  // scatteringTerm at the node (nonzero only)
    std::vector<double> weight;  // Simpson weight as the reference applies it
    std::vector<int> kind;       // 0: *weight, 1: *weight*4.0f, 2: *weight*2.0f
};

NodeSet make_nodes(const WeightingParameters &wp, float ds)
{
    NodeSet ns;
    const uint32_t n = wp.numStepsInt;
    const double bds = (float)(wp.scatter * ds);
    const double h = (float)((wp.maxBound - wp.minBound) / (float)n);
    const double weight = (double)(1.0f / 3.0f) * h;
    const double start = wp.minBound;
    const double eps = wp.mu, mu = wp.eps;  // the reference's argument swap
    auto add = [&](double p, int kind) {
        const double phase = GaussianPhase((float)p, (float)mu);
        const double f = p * std::exp(bds * phase - 1.0 * (eps * eps * p * p * 0.5));
        if (f == 0.0) return;
        ns.p.push_back(p);
        ns.factor.push_back(f);
        ns.weight.push_back(weight);
        ns.kind.push_back(kind);
    };
    add(start, 0);
    add((double)wp.maxBound, 0);
    for (uint32_t i = 1; i < n / 2; ++i) add(start + ((double)((2 * i) - 1) * h), 1);
    for (uint32_t i = 1; i < (n / 2) - 1; ++i) add(start + ((double)(2 * i) * h), 2);
    return ns;
}

float integrate_with_nodes(const NodeSet &ns, float curvature, float ds, const WeightingParameters &wp,
      std::vector<double> &scratch)
{
    const double kds = (float)(curvature * ds);
    scratch.resize(ns.p.size());
    for (size_t k = 0; k < ns.p.size(); ++k) {
        const double p = ns.p[k];
        const double sinTerm = (kds > 0.0) ? std::sin(kds * p) / kds : p;
        double v = (ns.factor[k] * sinTerm) * ns.weight[k];
        if (ns.kind[k] == 1) v = v * 4.0f;
        else if (ns.kind[k] == 2) v = v * 2.0f;
        scratch[k] = v;
    }
    std::sort(scratch.begin(), scratch.end());
    double sum = 0.0;
    for (const double v : scratch) sum += v;
    return (float)((sum / (2.0 * (double)TwistyPi * (double)TwistyPi)) * (double)wp.bias);
}
}  // namespace

float WeightLookupTableIntegral::Integrate(float curvature, const WeightingParameters &wp, float ds,
      std::vector<double> &scratch)
{
    const NodeSet ns = make_nodes(wp, ds);
    return integrate_with_nodes(ns, curvature, ds, wp, scratch);
}

// d T[i] / d eps, with m = eps the phase width and c = mu the regulariser
// (swapped as the reference passes them).  Per node p_j = start + j h with
// Simpson coefficient a_j (1/3, 4/3 or 2/3 of h):
//   T = B sum_j a_j h F(p_j),  F(p) = p e^{g(p)} s(p),
//   g = bds Phi(p; m) - c^2 p^2 / 2,  s = sin(kds p) / kds (p at kds = 0),
//   Phi = A(m) e^{-m p^2/2},  A = sqrt(pi m / 2) / (1 - e^{-2/m}),
// so with h = maxBound / n and maxBound = 10 / m (dh/dm = -h/m, dp_j/dm = -p_j/m):
//   dT/dm = B sum_j a_j h [ -F/m - (p_j/m) F_p + F_m ],
//   F_p = e^g (1 + p g_p) s + p e^g cos(kds p),  g_p = -bds m p Phi - c^2 p,
//   F_m = F bds Phi (1/(2m) - (2/m^2) e^{-2/m} / (1 - e^{-2/m}) - p^2/2).
// The reference evaluates Phi in float and rounds T to float; the derivative is
// that of the exact expression.  Entries the reference replaces by the table
// minimum (<= 0) take the derivative of that minimum's entry.
std::vector<double> WeightLookupTableIntegral::EpsDerivative(const WeightingParameters &wp, float ds,
      bool maxBoundFollowsEps, int threads)
{
    if (wp.numStepsInt < 4 || wp.numStepsInt % 2 != 0) throw std::runtime_error("numStepsInt must be even and >= 4");
    const uint32_t n = wp.numStepsInt, nb = wp.numCurvatureSteps;
    const double m = wp.eps, c = wp.mu;
    const double bds = (float)(wp.scatter * ds);
    const double h = (double)(wp.maxBound - wp.minBound) / (double)n;
    const double start = wp.minBound;
    const double B = (double)wp.bias / (2.0 * (double)TwistyPi * (double)TwistyPi);
    const double em = std::exp(-2.0 / m);
    const double A = std::sqrt(M_PI * m * 0.5) / (1.0 - em);
    const double dlnA = 0.5 / m - (2.0 / (m * m)) * em / (1.0 - em);
    const double follow = maxBoundFollowsEps ? 1.0 : 0.0;
    struct Node {
        double p, coef, eg, gp, phim;  // abscissa, a_j h, e^g, g_p, bds * dPhi/dm
    };
    std::vector<Node> nodes;
    auto add = [&](double p, double a) {
        const double phi = A * std::exp(-m * p * p * 0.5);
        const double g = bds * phi - c * c * p * p * 0.5;
        const double eg = std::exp(g);
        if (eg == 0.0) return;
        nodes.push_back({ p, a * h, eg, -bds * m * p * phi - c * c * p, bds * phi * (dlnA - p * p * 0.5) });
    };
    add(start, 1.0 / 3.0);
    add(start + (double)n * h, 1.0 / 3.0);
    for (uint32_t i = 1; i < n / 2; ++i) add(start + (double)(2 * i - 1) * h, 4.0 / 3.0);
    for (uint32_t i = 1; i < n / 2 - 1; ++i) add(start + (double)(2 * i) * h, 2.0 / 3.0);
    const float maxK = (float)((2.0 / (double)ds) * (double)1.01f);
    const float step = maxK / (float)(nb - 1);
    std::vector<double> d(nb), val(nb);
    if (threads < 1) threads = hardware_threads();
    std::vector<std::thread> pool;
    const uint32_t per = (nb + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint32_t b = std::min<uint32_t>(nb, t * per), e = std::min<uint32_t>(nb, b + per);
        pool.emplace_back([&, b, e] {
            for (uint32_t i = b; i < e; ++i) {
                const float k = (i == 0) ? 0.0f : ((float)(int64_t)i * step);
                const double kds = (float)(k * ds);
                double acc = 0.0, v = 0.0;
                for (const Node &nd : nodes) {
                    const double p = nd.p;
                    const double s = kds > 0.0 ? std::sin(kds * p) / kds : p;
                    const double sp = kds > 0.0 ? std::cos(kds * p) : 1.0;
                    const double F = p * nd.eg * s;
                    const double Fp = nd.eg * (1.0 + p * nd.gp) * s + p * nd.eg * sp;
                    const double Fm = F * nd.phim;
                    acc += nd.coef * (Fm - follow * (F + p * Fp) / m);
                    v += nd.coef * F;
                }
                d[i] = B * acc;
                val[i] = B * v;
            }
        });
    }
    for (auto &t : pool) t.join();
    uint32_t argmin = 0;
    for (uint32_t i = 1; i < nb; ++i)
        if ((float)val[i] < (float)val[argmin] && (float)val[i] >= 0.0f) argmin = i;
    for (uint32_t i = 0; i < nb; ++i)
        if ((float)val[i] <= 0.0f) d[i] = d[argmin];
    return d;
}

void WeightLookupTableIntegral::Initialize() { Initialize(hardware_threads()); }

void WeightLookupTableIntegral::Initialize(int threads)
{
    const WeightingParameters &wp = m_weightingParams;
    if (wp.numStepsInt < 4 || wp.numStepsInt % 2 != 0)
        throw std::runtime_error("numStepsInt must be even and >= 4");  // PathWeightUtils.cpp:382
    const uint32_t nb = wp.numCurvatureSteps;
    const NodeSet ns = make_nodes(wp, m_ds);
    std::vector<float> raw(nb);
    if (threads < 1) threads = hardware_threads();
    std::vector<std::thread> pool;
    const uint32_t per = (nb + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint32_t b = std::min<uint32_t>(nb, t * per), e = std::min<uint32_t>(nb, b + per);
        pool.emplace_back([&, b, e] {
            std::vector<double> scratch;
            for (uint32_t i = b; i < e; ++i) {
                const float k = (i == 0) ? 0.0f : (m_minCurvature + ((float)(int64_t)i * m_curvatureStepSize));
                raw[i] = integrate_with_nodes(ns, k, m_ds, wp, scratch);
            }
        });
    }
    for (auto &t : pool) t.join();
    // PathWeightUtils.cpp:303-347: min/max over non-negative entries (seeded
    // with entry 0), then entries <= 0 replaced by the minimum.
    float mn = raw[0], mx = raw[0];
    for (uint32_t i = 1; i < nb; ++i) {
        if (raw[i] < mn && raw[i] >= 0.0f) mn = raw[i];
        if (raw[i] > mx && raw[i] >= 0.0f) mx = raw[i];
    }
    for (float &v : raw)
        if (v <= 0.0) v = mn;
    m_lookupTable = std::move(raw);
    m_minSegmentWeight = mn;
    m_maxSegmentWeight = mx;
}

CachedMultiArclengthWeightLookupTable::CachedMultiArclengthWeightLookupTable(const WeightingParameters &wp,
      float minDs, float maxDs, uint32_t numDsSteps, const std::string &cacheRoot, int threads)
    : m_wp(wp)
    , m_minDs(minDs)
    , m_maxDs(maxDs)
    , m_numDsSteps(numDsSteps)
{
    if (numDsSteps < 2 || !(maxDs > minDs) || !(minDs > 0.0f))
        throw std::runtime_error("cached weight tables need 0 < minDs < maxDs and numDsSteps >= 2");
    std::ostringstream u;
    u << "minDs_" << hex_of(minDs) << "maxDs_" << hex_of(maxDs) << "numDsSteps_" << numDsSteps << "wp_"
      << wp.GenerateStringUUID().first;
    m_uuid = { u.str(), fnv1a(u.str()) };
    const std::string root = cacheRoot.empty() ? std::string(".") : cacheRoot;
    const std::string dir = root + "/CachedWeightTables";
    m_cachePath = dir + "/" + std::to_string(m_uuid.second) + ".cwt";
    auto make = [&](uint32_t i) -> std::unique_ptr<BaseWeightLookupTable> {
        if (wp.weightingMethod == WeightingMethod::RadiativeTransfer)
            return std::make_unique<WeightLookupTableIntegral>(wp, DsAt(i));
        return std::make_unique<SimpleWeightLookupTable>(wp, DsAt(i));
    };
    std::ifstream in(m_cachePath, std::ios::binary);
    if (in.is_open()) {
        for (uint32_t i = 0; i < numDsSteps; ++i) {
            m_tables.push_back(make(i));
            m_tables.back()->InitializeFromStream(in);
        }
        m_loaded = true;
        return;
    }
    for (uint32_t i = 0; i < numDsSteps; ++i) {
        m_tables.push_back(make(i));
        if (auto *rt = dynamic_cast<WeightLookupTableIntegral *>(m_tables.back().get())) rt->Initialize(threads);
        else m_tables.back()->Initialize();
    }
    std::error_code ec;
    std::filesystem::create_directories(dir + "/" + std::to_string(m_uuid.second), ec);
    std::ofstream out(m_cachePath, std::ios::binary);
    if (!out.is_open()) throw std::runtime_error("Failed to export weight table: " + m_cachePath);
    for (const auto &t : m_tables) t->StreamOut(out);
    out.close();
    for (const auto &t : m_tables) t->ExportValues(dir + "/" + std::to_string(m_uuid.second));
}

float CachedMultiArclengthWeightLookupTable::DsAt(uint32_t i) const
{
    const float step = (m_maxDs - m_minDs) / (float)(m_numDsSteps - 1);
    return m_minDs + (float)i * step;
}

size_t CachedMultiArclengthWeightLookupTable::ClosestIndex(float ds) const
{
    size_t best = 0;
    float bestDist = std::numeric_limits<float>::max();
    for (size_t i = 0; i < m_tables.size(); ++i) {
        const float d = std::abs(m_tables[i]->GetDs() - ds);
        if (d < bestDist) {
            bestDist = d;
            best = i;
        }
    }
    return best;
}

const BaseWeightLookupTable *CachedMultiArclengthWeightLookupTable::GetWeightLookupTable(float ds) const
{
    return m_tables.empty() ? nullptr : m_tables[ClosestIndex(ds)].get();
}

// PathWeightUtils.cpp:436-485.
SimpleWeightLookupTable::SimpleWeightLookupTable(const WeightingParameters &wp, float ds)
    : BaseWeightLookupTable(wp, ds, -1.0f, 1.0f)
{
}

void SimpleWeightLookupTable::Initialize()
{
    const WeightingParameters &wp = m_weightingParams;
    auto value = [&](float curvature) -> float {
        const float alpha = 1.0 / (wp.scatter * m_ds * wp.mu);
        const float leftComponent = std::exp(-1.0 * alpha) / (m_ds * m_ds * m_ds);
        const float rightComponent = std::exp(alpha * curvature);
        return leftComponent * rightComponent * wp.bias;
    };
    m_lookupTable[0] = value(m_minCurvature);
    float mn = m_lookupTable[0], mx = m_lookupTable[0];
    for (uint32_t i = 1; i < wp.numCurvatureSteps; ++i) {
        const float v = value(m_minCurvature + (float)i * m_curvatureStepSize);
        m_lookupTable[i] = v;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    m_minSegmentWeight = mn;
    m_maxSegmentWeight = mx;
}

// ---------------------------------------------------------------------------
// Path normaliser, PathWeightUtils.cpp:643-752, accumulated in long double
// (the reference uses Boost cpp_dec_float_100).
// ---------------------------------------------------------------------------
namespace {
long double K4(float z, float ds)
{
    float result = (z < 2.0) ? 1.0f : 0.0f;
    result /= z;
    result *= 2.0 * TwistyPi;
    const float invDs = (1.0 / ds);
    result *= (invDs * invDs * invDs);
    return result;
}

float rmag(float x, float z) { return std::sqrt((z * z) + 1.0 - (2.0 * z * x)); }

// F(6,5) (PathWeightUtils.cpp:659-687, numSteps 1000 there): left Riemann
// rule over F5 = K4(rmag), i.e. K4 convolved with the unit sphere.
long double F6(float x, float z, float ds, uint32_t steps)
{
    const float r = rmag(x, z);
    const float dx = (1.0f - -1.0f) / steps;
    long double acc = 0.0;
    for (uint32_t i = 0; i < steps; ++i) acc += K4(rmag(-1.0f + i * dx, r), ds);
    return acc * ((2.0 * TwistyPi) * dx);
}

// F(7,6): the same rule over F6.
long double F7(float x, float z, float ds, uint32_t steps)
{
    const float r = rmag(x, z);
    const float dx = (1.0f - -1.0f) / steps;
    long double acc = 0.0;
    for (uint32_t i = 0; i < steps; ++i) acc += F6(-1.0f + i * dx, r, ds, steps);
    return acc * ((2.0 * TwistyPi) * dx);
}

// K(N) macro (PathWeightUtils.cpp:689-711, numSteps 10000 there): the rule
// over F7 for every order, so K5 == K6.  The outer nodes are dealt to the
// host threads; their partial sums are added in thread order.
long double KMacro(float z, float ds, uint32_t outerSteps, uint32_t innerSteps)
{
    const float dx = (1.0f - -1.0f) / outerSteps;
    const int threads = std::max(1, std::min<int>(hardware_threads(), (int)outerSteps));
    std::vector<long double> part(threads, 0.0L);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint32_t i = t; i < outerSteps; i += threads) part[t] += F7(-1.0f + i * dx, z, ds, innerSteps);
        });
    for (auto &th : pool) th.join();
    long double acc = 0.0;
    for (long double v : part) acc += v;
    return acc * ((2.0 * TwistyPi) * dx);
}
}  // namespace

double NormLog10(int numberOfSegments, float ds, const PerturbUtils::BoundaryConditions &bc, uint32_t outerSteps,
      uint32_t innerSteps)
{
    const Vec3 zVec = (bc.m_endPos - bc.m_startPos) * (1.0f / ds) - bc.m_endDir - bc.m_startDir;
    const float z = zVec.Magnitude();
    long double v = 1.0L;
    if (numberOfSegments == 4) v = K4(z, ds);
    else if (numberOfSegments == 5 || numberOfSegments == 6) v = KMacro(z, ds, outerSteps, innerSteps);
    return (double)std::log10(v);
}

}  // namespace PathWeighting
}  // namespace twisty
