// generate.hip -- independent-path estimator on gfx950: the paths of
// ExperimentBase::MSegmentPathGenerationMC (Experiments/ExperimentBase.cpp:
// 1246-1386), each built from scratch by recursive sphere intersection and
// weighted like the perturbed paths.
//
// A group of G lanes owns one path; its N + 1 points live in LDS in the
// [row][component][lane] layout of the forward kernel (point i in row i / G,
// column i mod G).  The recursion is run level by level: level l has 2^l
// independent centre placements (ResolveEvenNumberOfSegments, :806-983), the
// last level the two- or three-segment leaves (ResolveTwoSegments :617-698,
// ResolveThreeSegments :702-804); the lanes of the group take the nodes of a
// level in turn.  The points are then turned into segment vectors in place and
// weighted by weigh_state (tangents, curvatures, table lookup, as the forward
// pass), and the weights go into the same LSE partials.
//
// Draws: node n (heap numbering) of path p reads one Philox4x32-10 block with
// counter (n, 0x80000000 | sub, p lo, p hi) and the seed as key; sub = 1 is the
// two-segment step inside a three-segment leaf.  phi is drawn as
// cos(phi) = 1 - 2 u vmax, the reference's acos(1 - 2 U(0, vmax)) without the
// acos, and square roots are taken in factored form where the reference's
// difference of squares cancels.  A step the reference aborts with an
// exception (spheres that do not meet, a NaN phi extent) leaves NaN points and
// the path is counted invalid.
#include "generate.hpp"
#include "path_kernels.hpp"

namespace twisty_hip {

namespace {

struct F3 {
    float x, y, z;
};

// Reciprocal, reciprocal square root, quotient and square root at the
// hardware instructions' precision (about 1 ulp): the sphere-intersection
// geometry tolerates it, and the IEEE-correct division and square root take
// several times the instructions.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float hsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ F3 f3(float x, float y, float z) { return { x, y, z }; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ F3 cross3(F3 a, F3 b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ F3 unit(F3 a) { return a * __builtin_amdgcn_rsqf(dot3(a, a)); }
__device__ __forceinline__ F3 nan3() { return { __int_as_float(0x7fc00000), __int_as_float(0x7fc00000), __int_as_float(0x7fc00000) }; }

struct Words {
    float u0, u1, u3;
    uint32_t coin;
};

__device__ __forceinline__ Words gen_words(uint64_t seed, uint64_t path, uint32_t node, uint32_t sub)
{
    const U4 r = philox4x32_10(node, 0x80000000u | sub, (uint32_t)path, (uint32_t)(path >> 32), (uint32_t)seed,
          (uint32_t)(seed >> 32));
    return { (float)(r.x >> 8) * 0x1p-24f, (float)(r.y >> 8) * 0x1p-24f, (float)(r.w >> 8) * 0x1p-24f, r.z & 1u };
}

// x = z cross e, y = z cross x (ExperimentBase.cpp:868-875), e = +x unless
// |z . x| > limit, then +y.
__device__ __forceinline__ void frame(F3 z, float limit, F3 &x, F3 &y)
{
    const F3 other = fabsf(z.x) > limit ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
    x = unit(cross3(z, other));
    y = unit(cross3(z, x));
}

// cos and sin of phi with cos(phi) = 1 - 2 u vmax
__device__ __forceinline__ void phi_of(float u, float vmax, float &c, float &s)
{
    const float v = u * vmax;
    c = fmaf(-2.0f, v, 1.0f);
    s = 2.0f * hsqrt(fmaxf(0.0f, v * (1.0f - v)));
}

__device__ __forceinline__ F3 spherical(F3 x, F3 y, F3 z, float c, float s, float ct, float st)
{
    return x * (s * ct) + y * (s * st) + z * c;
}

__device__ F3 resolve_two(F3 L, F3 R, float ds, const Words &g)
{
    const F3 dv = R - L;
    const float d = hsqrt(dot3(dv, dv));
    if (fabsf(2.0f * ds - d) < 0.001f) return (L + R) * 0.5f;
    if (2.0f * ds < d) return nan3();
    float st, ct;
    sincospi_turn(g.u1, st, ct);
    if (d < 0.001f) {
        float c, s;
        phi_of(g.u0, 1.0f, c, s);
        return L + f3(s * ct, s * st, c) * ds;
    }
    const F3 n = dv * rcp(d);
    const F3 other = fabsf(n.x) >= 0.99f ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
    const F3 m = unit(cross3(n, other));
    const float hd = 0.5f * d;
    const float h = ds > hd ? hsqrt((ds - hd) * (ds + hd)) : 0.0f;
    return (L + R) * 0.5f + (m * ct + cross3(n, m) * st) * h;
}

__device__ F3 resolve_three_first(F3 L, F3 R, float ds, const Words &g)
{
    const F3 dv = R - L;
    const float d = hsqrt(dot3(dv, dv));
    if (fabsf(3.0f * ds - d) < 0.001f) return L + dv * (1.0f / 3.0f);
    if (3.0f * ds < d) return nan3();
    float st, ct;
    sincospi_turn(g.u1, st, ct);
    if (d < 0.001f) {
        float c, s;
        phi_of(g.u0, 1.0f, c, s);
        return L + f3(s * ct, s * st, c) * ds;
    }
    const F3 z = dv * rcp(d);
    F3 x, y;
    frame(z, 0.999f, x, y);
    const float r1 = ds, r2 = 2.0f * ds;
    float vmax;
    if (d + r1 < r2) {
        vmax = 1.0f;
    } else {
        const float h = 0.5f + qdiv(r1 * r1 - r2 * r2, 2.0f * d * d);
        const float t = fmaf(-h * d, h * d, r1 * r1);
        float a;
        if (fabsf(t) < 0.001f) a = 0.0f;
        else if (t < 0.0f) return nan3();
        else a = hsqrt(t);
        const float sE = fminf(1.0f, qdiv(a, r1));
        const float cE = hsqrt((1.0f - sE) * (1.0f + sE));
        vmax = (h * d < 0.0f) ? 0.5f + 0.5f * cE : qdiv(0.5f * sE * sE, 1.0f + cE);
    }
    float c, s;
    phi_of(g.u0, vmax, c, s);
    return L + unit(spherical(x, y, z, c, s, ct, st)) * ds;
}

__device__ F3 resolve_even_centre(uint32_t n, F3 L, F3 R, float ds, const Words &g)
{
    const F3 dv = R - L;
    const float d = hsqrt(dot3(dv, dv));
    if (fabsf((float)n * ds - d) < 0.001f) return (L + R) * 0.5f;
    const float rps = ds * (float)(n / 2);
    if (2.0f * rps < d) return nan3();
    float st, ct;
    sincospi_turn(g.u1, st, ct);
    const F3 z = unit(dv);
    F3 x, y;
    frame(z, 0.999f, x, y);
    if (d < 0.001f) {
        float c, s;
        phi_of(g.u0, 1.0f, c, s);
        return L + spherical(x, y, z, c, s, ct, st) * (rps * cbrtf(g.u3));
    }
    const float vmax = 0.5f * (1.0f - qdiv(d, 2.0f * rps));
    float c, s;
    phi_of(g.u0, vmax, c, s);
    const float minPct = qdiv(qdiv(0.5f * d, c), rps);
    const float radius = rps * cbrtf(fmaf(1.0f - minPct, g.u3, minPct));
    if (g.coin == 0u) return L + spherical(x, y, z, c, s, ct, st) * radius;
    return R + spherical(x, y, z, -c, s, ct, st) * radius;
}

__device__ __forceinline__ void atomic_min_f32(float *addr, float v)
{
    unsigned int *a = reinterpret_cast<unsigned int *>(addr);
    unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (v < __uint_as_float(old)) {
        const unsigned int prev = atomicCAS(a, old, __float_as_uint(v));
        if (prev == old) break;
        old = prev;
    }
}

}  // namespace

// Nearest table to ds in the multi-ds set: the tables' ds are a uniform grid,
// so the nearest index is the rounded one or a neighbour; the three are
// compared in increasing order with a strict < (first table on ties).
__device__ __forceinline__ uint32_t nearest_table(const GenParams &p, float ds)
{
    const uint32_t n = p.numTables;
    if (n == 1) return 0;
    const float d0 = p.multiMeta[0], d1 = p.multiMeta[5 * (n - 1)];
    const float pos = (ds - d0) / (d1 - d0) * (float)(n - 1);
    int i0 = (int)rintf(pos);
    i0 = min(max(i0, 0), (int)n - 1);
    uint32_t best = 0;
    float bestDist = INFINITY;
    for (int i = max(i0 - 1, 0); i <= min(i0 + 1, (int)n - 1); ++i) {
        const float dist = fabsf(p.multiMeta[5 * i] - ds);
        if (dist < bestDist) {
            bestDist = dist;
            best = (uint32_t)i;
        }
    }
    return best;
}

// Build-defined natural logarithm of y in (0, 1] in FP64 from basic operations,
// the same steps as the oracle's, so both form the same bits: y = m 2^e with m
// in [sqrt(1/2), sqrt(2)), ln m = 2 s (1 + z/3 + ... + z^11/23), s = (m-1)/(m+1),
// z = s^2.
__device__ __forceinline__ double bd_log(double y)
{
    int e = 0;
    double m = frexp(y, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = __ddiv_rn(m - 1.0, m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
#pragma unroll
    for (int k = 10; k >= 0; --k) p = fma(p, z, 1.0 / (double)(2 * k + 1));
    return fma((double)e, 0.69314718055994530942, (2.0 * s) * p);
}

// MODE: kGenSingle one table and ds; kGenMulti a ds per path, drawn uniform,
// and the nearest cached table; kGenChains the two-table overload's
// Metropolis-Hastings chains (ExperimentBase.cpp:1599-1818): per proposal a
// truncated-exponential ds and the nearest cached table, the sampling weight,
// and the acceptance test against the chain's current weight.
template<int G, int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK) generate_kernel(const GenParams p)
{
    constexpr bool MULTI = MODE != kGenSingle;
    constexpr uint32_t SLOT = 3u * BLOCK;
    constexpr uint32_t groupsPerBlock = BLOCK / G;
    constexpr int SMAX = 16;  // segment rows per lane: ceil(N / G) <= ceil((N + 1) / G) <= 16
    extern __shared__ float smem[];
    const uint32_t N = p.N;
    const uint32_t tabOff = p.rows * SLOT;
    if constexpr (!MULTI) {
        for (int i = threadIdx.x; i < p.tableSize; i += BLOCK) smem[tabOff + i] = p.table[i];
        __syncthreads();
    }
    TableView tv = make_table_view(smem + tabOff, p.tableSize, p.minK, p.maxK, p.step);
    const uint32_t lane = threadIdx.x;
    const uint32_t j = lane & (G - 1);
    const uint32_t groupBase = lane - j;
    const uint32_t S = (N + G - 1) / G;
    float ds = p.ds, invDs = p.invDs, absorbLog2 = 0.0f;
    const F3 start = f3(p.start[0], p.start[1], p.start[2]);
    const F3 end = f3(p.end[0], p.end[1], p.end[2]);
    const F3 startDir = f3(p.startDir[0], p.startDir[1], p.startDir[2]);
    const F3 endDir = f3(p.endDir[0], p.endDir[1], p.endDir[2]);
    auto at = [&](uint32_t i) { return (i / G) * SLOT + groupBase + (i & (uint32_t)(G - 1)); };
    auto ld = [&](uint32_t i) {
        const uint32_t a = at(i);
        return f3(smem[a], smem[a + BLOCK], smem[a + 2 * BLOCK]);
    };
    auto st = [&](uint32_t i, F3 v) {
        const uint32_t a = at(i);
        smem[a] = v.x;
        smem[a + BLOCK] = v.y;
        smem[a + 2 * BLOCK] = v.z;
    };

    LseAcc acc;
    lse_init(acc);
    uint32_t clamped = 0;
    float minL = INFINITY;
    const uint32_t n0 = N - 2;
    // Builds and weighs path `path` (output slot `local`); every lane of the
    // group returns the group's totals.
    auto one_path = [&](uint64_t local, uint64_t path) -> PathLog {
        if constexpr (MULTI) {
            const Words w0 = gen_words(p.seed, path, 0, 0);
            if constexpr (MODE == kGenChains) {
                // ds = minDs + x, x exponential of rate t truncated to [0, maxDs], by the
                // inverse distribution function (the reference rejects draws past maxDs +
                // minDs, ExperimentBase.cpp:1662-1667: the same distribution)
                const double y = fma(-(double)w0.u0, p.oneMinusQ, 1.0);
                ds = (float)((double)p.dsLo + __ddiv_rn(-bd_log(y), (double)p.extinction));
            } else {
                // ds = U(0,1) (maxDs - minDs) + minDs (ExperimentBase.cpp:1486-1487)
                ds = fmaf(w0.u0, p.dsHi - p.dsLo, p.dsLo);
            }
            // the nearest cached table, curvature with arclength ds N (:1512-1518)
            const uint32_t t = nearest_table(p, ds);
            const float *m = p.multiMeta + 5 * t;
            tv = make_table_view(p.multiTables + (uint64_t)t * p.tableStride, p.tableSize, m[1], m[2], m[3]);
            absorbLog2 = m[4];
            invDs = (float)N / (ds * (float)N);
        }
        if (j == 0) {
            st(0, start);
            st(1, start + startDir * ds);  // hinge points, ExperimentBase.cpp:1255-1261
            st(N - 1, end - endDir * ds);
            st(N, end);
        }
        wave_sync_lds();
        for (uint32_t lv = 0; lv < p.levels; ++lv) {
            const uint32_t nodes = 1u << lv, span = n0 >> lv;
            for (uint32_t q = j; q < nodes; q += G) {
                const uint32_t l = 1 + q * span;
                st(l + span / 2, resolve_even_centre(span, ld(l), ld(l + span), ds, gen_words(p.seed, path, nodes + q, 0)));
            }
            wave_sync_lds();
        }
        {
            const uint32_t nodes = 1u << p.levels, span = p.leafSpan;
            for (uint32_t q = j; q < nodes; q += G) {
                const uint32_t l = 1 + q * span;
                if (span == 2) {
                    st(l + 1, resolve_two(ld(l), ld(l + 2), ds, gen_words(p.seed, path, nodes + q, 0)));
                } else {
                    const F3 R = ld(l + 3);
                    const F3 p1 = resolve_three_first(ld(l), R, ds, gen_words(p.seed, path, nodes + q, 0));
                    st(l + 1, p1);
                    st(l + 2, resolve_two(p1, R, ds, gen_words(p.seed, path, nodes + q, 1)));
                }
            }
            wave_sync_lds();
        }
        // points -> segment vectors in place (segment i = p_{i+1} - p_i sits where p_i was)
        float sx[SMAX], sy[SMAX], sz[SMAX];
#pragma unroll
        for (int s = 0; s < SMAX; ++s) {
            const uint32_t i = (uint32_t)s * G + j;
            if ((uint32_t)s < S && i < N) {
                const F3 d = ld(i + 1) - ld(i);
                sx[s] = d.x;
                sy[s] = d.y;
                sz[s] = d.z;
            }
        }
        wave_sync_lds();
#pragma unroll
        for (int s = 0; s < SMAX; ++s) {
            const uint32_t i = (uint32_t)s * G + j;
            if ((uint32_t)s < S && i < N) {
                smem[s * SLOT + lane] = sx[s];
                smem[s * SLOT + BLOCK + lane] = sy[s];
                smem[s * SLOT + 2 * BLOCK + lane] = sz[s];
            }
        }
        wave_sync_lds();
        PathLog pl;
        if (!p.perJointCurv) {
            pl = weigh_state<G, BLOCK>(smem, lane, j, groupBase, N, invDs, tv, clamped);
        } else {
            pl = weigh_state<G, BLOCK>(smem, lane, j, groupBase, N, invDs, tv, clamped,
                  [&](uint32_t i, const JointEval &je) { p.perJointCurv[local * (uint64_t)(N - 1) + i] = je.kappa; });
        }
        if constexpr (MULTI) pl.lsum = ff_add_f(pl.lsum, absorbLog2);  // the table's own absorption
        if constexpr (MODE == kGenChains) {
            // log2 of the sampling weight 1 / (t exp(-t (ds - minDs) N)), formed in
            // float as ExperimentBase.cpp:1669-1670 forms it
            const float t = p.extinction;
            const float sw = 1.0f / (t * expf(-t * (ds - p.dsLo) * (float)N));
            const double l2 = log2((double)sw);
            const float hi = (float)l2;
            pl.lsum = ff_add_f(ff_add_f(pl.lsum, hi), (float)(l2 - (double)hi));
        }
        wave_sync_lds();  // joints read other lanes' columns before the next path's points
        return pl;
    };
    // A path's log10 weight as it is reported (NaN when invalid).
    auto log10_of = [&](const PathLog &pl, bool &invalid) -> float {
        invalid = (pl.flags & 1u) || ((pl.flags & 6u) == 6u);
        if (invalid) return __int_as_float(0x7fc00000);
        if (pl.flags & 2u) return -INFINITY;
        if (pl.flags & 4u) return INFINITY;
        return fmaf(pl.lsum.hi, p.log10of2, fmaf(pl.lsum.lo, p.log10of2, p.absorbLog10PerPath));
    };
    // Adds a valid path (or the chain state) of log10 weight v and log2 sum ls.
    auto add_valid = [&](float v, FF ls) {
        ++acc.nValid;
        if (v == INFINITY) lse_add(acc, INFINITY, 0.0f);
        else if (v != -INFINITY) lse_add(acc, ls.hi, ls.lo);
        minL = fminf(minL, v);
    };

    if constexpr (MODE != kGenChains) {
        for (uint64_t local = (uint64_t)blockIdx.x * groupsPerBlock + lane / G; local < p.numPaths;
              local += (uint64_t)gridDim.x * groupsPerBlock) {
            const PathLog pl = one_path(local, p.pathOffset + local);
            if (j == 0) {
                bool invalid;
                const float outv = log10_of(pl, invalid);
                if (invalid) ++acc.nInvalid;
                else add_valid(outv, pl.lsum);
                if (p.perPathLog10) p.perPathLog10[local] = outv;
            }
        }
    } else {
        // One chain per group at a time, its proposals in order.  Every lane of
        // the group holds the same totals, so the acceptance test is the same in
        // all of them.
        for (uint64_t chain = (uint64_t)blockIdx.x * groupsPerBlock + lane / G; chain < p.numChains;
              chain += (uint64_t)gridDim.x * groupsPerBlock) {
            bool first = true;
            float curV = __int_as_float(0x7fc00000);  // the chain's current log10 weight
            FF curS { 0.0f, 0.0f };
            double curN = 0.0;                        // and with the normaliser, in FP64
            for (uint64_t st = 0; st < p.chainLength; ++st) {
                const uint64_t local = chain * p.chainLength + st;
                const uint64_t path = (p.pathOffset + chain) * p.chainLength + st;
                const PathLog pl = one_path(local, path);
                bool invalid;
                const float outv = log10_of(pl, invalid);
                if (invalid) {
                    // skipped: the chain keeps its state (ExperimentBase.cpp:1720-1723)
                    if (j == 0) ++acc.nInvalid;
                } else {
                    // the FP64 weight with the normaliser; pow(10, w) as the reference forms it
                    const double w = (outv == INFINITY || outv == -INFINITY)
                          ? (double)outv
                          : ((double)pl.lsum.hi + (double)pl.lsum.lo) * 0.30102999566398119521 + p.normLog10;
                    bool take = first;
                    if (!first) {
                        const double ratio = exp10(w) / exp10(curN);
                        take = (double)gen_words(p.seed, path, 0, 0).u1 <= ratio;  // :1738-1755
                    }
                    if (take) {
                        curV = outv;
                        curS = pl.lsum;
                        curN = w;
                    }
                    first = false;
                    if (j == 0) add_valid(curV, curS);  // :1757-1784
                }
                if (j == 0) {
                    if (p.perPathLog10) p.perPathLog10[local] = outv;
                    if (p.chainLog10) p.chainLog10[local] = curV;
                }
            }
        }
    }
    acc.nClamped = clamped;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) minL = fminf(minL, __shfl_xor(minL, o, 64));
    if ((lane & 63u) == 0 && minL < INFINITY) atomic_min_f32(p.minLog10, minL);
    block_reduce_lse(acc, smem, p.partials + blockIdx.x);
}

#define TWISTY_GEN_INSTANCES(X) X(2, 256) X(4, 256) X(8, 256) X(16, 256) X(32, 256) X(64, 256)

template<int G, int B, int M>
static hipError_t launch_gen_gb(const GenParams &p, uint32_t grid, uint32_t lds, hipStream_t s)
{
    auto *fn = generate_kernel<G, B, M>;
    hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, s, p);
    return hipGetLastError();
}

hipError_t launch_generate(const GenParams &p, uint32_t G, int mode, uint32_t grid, uint32_t block, uint32_t lds,
      hipStream_t s)
{
#define X(g, b)                                                                                                   \
    if (G == g && block == b) {                                                                                   \
        if (mode == kGenChains) return launch_gen_gb<g, b, kGenChains>(p, grid, lds, s);                          \
        return mode == kGenMulti ? launch_gen_gb<g, b, kGenMulti>(p, grid, lds, s)                                \
                                 : launch_gen_gb<g, b, kGenSingle>(p, grid, lds, s);                              \
    }
    TWISTY_GEN_INSTANCES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace twisty_hip
