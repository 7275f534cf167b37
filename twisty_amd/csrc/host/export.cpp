// export.cpp -- path-batch export files of ExperimentRunner.cpp:210-262,303-396
// and FullExperimentRunnerOptimalPerturb.cpp:599-858 (outputPathBatches).
//
// Under <experimentDir>/GeneratedCurves/:
//   BoundaryConditions.bcf        4 x float3 (start pos, start dir, end pos,
//                                 end dir), float arclength, uint32 N
//   <prepend>Paths_Metadata.pmd   one text line per batch: "thread batch count"
//   <prepend>Paths_Binary.pbd     per path (N+1) float3 points
//   <prepend>Log10PathWeights_Binary.bdt   per path one double, the weight
//                                 10^(log10 w + log10 normaliser) (the reference
//                                 stores the linear weight under this name)
//   <prepend>Log10PathWeights_Text.pwt     every 10000th weight of a batch, text
//   <prepend>FiveSegment_Binary.bdt        per path float3 (phi1, theta1, theta2)
//                                 of the five-segment parameterisation (N = 5;
//                                 zeros otherwise)
//   index.json                    {"experiment_name", "path_batch_links": []}
// Batches hold up to 1e6 exported (valid) paths, as the reference's cache.
#include <cmath>
#include <filesystem>
#include <fstream>
#include <stdexcept>

#include "twisty_host.hpp"

namespace twisty {

namespace {
constexpr uint64_t kBatchPaths = 1000000;  // ExportPathBatchCacheSize

void open_or_throw(std::ofstream &f, const std::string &path, bool binary)
{
    f.open(path, binary ? std::ios::binary : std::ios::out);
    if (!f.is_open()) throw std::runtime_error("Failed to open " + path);
}

Vec3 cross(const Vec3 &a, const Vec3 &b)
{
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// FullExperimentRunnerOptimalPerturb.cpp:756-793 (float arithmetic).
Vec3 five_segment_angles(const Vec3 *p)
{
    const Vec3 segment1Dir = (p[2] - p[1]).Normalized();
    const float phi1 = std::acos(segment1Dir.z);
    const float theta1 = std::atan2(segment1Dir.y, segment1Dir.x);
    const Vec3 z = (p[4] - p[2]).Normalized();
    Vec3 other(1.0f, 0.0f, 0.0f);
    if (std::fabs(z.Dot(other)) >= 0.99f) other = Vec3(0.0f, 1.0f, 0.0f);
    const Vec3 x = cross(z, other).Normalized();
    const Vec3 y = cross(z, x);
    const Vec3 v = (p[3] - p[2]).Normalized();
    return Vec3(phi1, theta1, std::atan2(v.Dot(y), v.Dot(x)));
}
}  // namespace

PathBatchWriter::PathBatchWriter(const std::string &experimentDir, const std::string &prepend,
      const std::string &experimentName, const PerturbUtils::BoundaryConditions &bc, uint32_t numSegments)
    : m_numSegments(numSegments)
{
    const std::filesystem::path dir = std::filesystem::path(experimentDir) / "GeneratedCurves";
    std::filesystem::create_directories(dir);
    m_dir = dir.string();
    m_experimentName = experimentName;
    {
        std::ofstream bcf;
        open_or_throw(bcf, (dir / "BoundaryConditions.bcf").string(), true);
        const Vec3 v[4] = { bc.m_startPos, bc.m_startDir, bc.m_endPos, bc.m_endDir };
        for (const Vec3 &x : v) bcf.write(reinterpret_cast<const char *>(&x), sizeof(Vec3));
        bcf.write(reinterpret_cast<const char *>(&bc.arclength), sizeof(float));
        bcf.write(reinterpret_cast<const char *>(&numSegments), sizeof(uint32_t));
    }
    open_or_throw(m_pmd, (dir / (prepend + "Paths_Metadata.pmd")).string(), false);
    open_or_throw(m_pbd, (dir / (prepend + "Paths_Binary.pbd")).string(), true);
    open_or_throw(m_bdt, (dir / (prepend + "Log10PathWeights_Binary.bdt")).string(), true);
    open_or_throw(m_pwt, (dir / (prepend + "Log10PathWeights_Text.pwt")).string(), false);
    open_or_throw(m_fiveFile, (dir / (prepend + "FiveSegment_Binary.bdt")).string(), true);
}

void PathBatchWriter::Add(const float *positions, const float *log10w, uint64_t numPaths, double normLog10)
{
    const uint32_t P = m_numSegments + 1;
    for (uint64_t q = 0; q < numPaths; ++q) {
        if (std::isnan(log10w[q])) continue;  // invalid paths are not exported
        const Vec3 *pts = reinterpret_cast<const Vec3 *>(positions + q * 3ull * P);
        m_points.insert(m_points.end(), pts, pts + P);
        m_weights.push_back(std::pow(10.0, (double)log10w[q] + normLog10));
        m_fiveCache.push_back(m_numSegments == 5 ? five_segment_angles(pts) : Vec3(0.0f, 0.0f, 0.0f));
        if (m_weights.size() == kBatchPaths) Flush();
    }
}

void PathBatchWriter::Flush()
{
    const uint64_t n = m_weights.size();
    if (n == 0) return;
    m_pmd << 0 << " " << m_batchIdx << " " << n << std::endl;
    m_pbd.write(reinterpret_cast<const char *>(m_points.data()), m_points.size() * sizeof(Vec3));
    m_bdt.write(reinterpret_cast<const char *>(m_weights.data()), n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i)
        if ((static_cast<uint32_t>(i) % 10000) == 0) m_pwt << m_weights[i] << std::endl;
    m_fiveFile.write(reinterpret_cast<const char *>(m_fiveCache.data()), n * sizeof(Vec3));
    m_points.clear();
    m_weights.clear();
    m_fiveCache.clear();
    ++m_batchIdx;
}

void PathBatchWriter::Finish()
{
    Flush();
    std::ofstream idx;
    open_or_throw(idx, m_dir + "/index.json", false);
    std::string name;
    for (char c : m_experimentName) {
        if (c == '"' || c == '\\') name += '\\';
        name += c;
    }
    idx << "{\n    \"experiment_name\": \"" << name << "\",\n    \"path_batch_links\": []\n}\n";
}

}  // namespace twisty
