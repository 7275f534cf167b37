// device_common.hpp -- device-side building blocks of the optimal-perturb hot path.
//
// Counter-based draws, the table lookup of PathWeighters.cpp:27-104 and the
// log-domain accumulator used by every kernel in this library.  Compiled for
// gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace twisty_hip {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter =
// (perturbation, stream, path_lo, path_hi), key = (seed_lo, seed_hi).  Stands in
// for the reference's per-thread std::mt19937_64 / curand XORWOW states
// (FullExperimentRunnerOptimalPerturb.cpp:113-120, ...Optimized_GPU.cu:678-686):
// no state, no init kernel, any path range reproducible from its counter.
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ k in one v_bitop3_b32 (truth table 0x96, the three-input XOR of
// gfx950); written as two XORs the round takes two instructions per word.
__device__ __forceinline__ uint32_t xor3_key(uint32_t a, uint32_t b, uint32_t k)
{
    return __builtin_amdgcn_bitop3_b32(a, b, k, 0x96);
}

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
      uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3_key((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3_key((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return { c0, c1, c2, c3 };
}

// One perturbation's draw: diff ~ U{2..maxDiff}, left ~ U{1..N-diff-1},
// half-angle = pi*x with x exactly representable in float
// (FullExperimentRunnerOptimalPerturb.cpp:483-502 ranges).
struct Draw {
    uint32_t diff;
    uint32_t left;
    float ch;  // cos(theta/2)
    float sh;  // sin(theta/2)
};

// A draw as the rotation kernels consume it: packed indices and the
// trigonometric terms of the rotation matrix (cos theta, sin theta,
// 1 - cos theta = 2 sin^2(theta/2) without cancellation).
struct DrawRec {
    uint32_t packed;  // diff | left << 16
    float c, s, C;
};

__device__ __forceinline__ DrawRec pack_draw(const Draw &d)
{
    DrawRec r;
    r.packed = d.diff | (d.left << 16);
    r.c = d.ch * d.ch - d.sh * d.sh;
    r.s = 2.0f * d.sh * d.ch;
    r.C = 2.0f * d.sh * d.sh;
    return r;
}

// sin(pi x) and cos(pi x) for |x| < 1/2, the range of the half-angle draw.
// For |x| > 1/4 the argument becomes y = 1/2 - |x| (exact in float) and sin
// and cos swap, so both polynomials only cover |y| <= 1/4: sin(pi y) =
// y P(y^2) with P of degree 3 and cos(pi y) = Q(y^2) with Q of degree 4, the
// coefficients a relative-weighted least-squares fit at 4000 Chebyshev nodes
// of [0, 1/4] rounded to float, evaluated by Horner's rule with FMA.  Over
// every x the draw produces (odd multiples of 2^-25) the error is at most 1.7
// ulp (mean 0.35 ulp) against FP64 sin/cos.  About 15 VALU instead of the
// general sincospif's ~40 with its full-range reduction.
__device__ __forceinline__ void sincospi_half(float x, float &s, float &c)
{
    const float a = fabsf(x);
    const bool swap = a > 0.25f;
    const float y = swap ? 0.5f - a : a;
    const float t = y * y;
    float p = -0x1.2db8e8p-1f;
    p = fmaf(p, t, 0x1.465f8ep+1f);
    p = fmaf(p, t, -0x1.4abbbcp+2f);
    p = fmaf(p, t, 0x1.921fb6p+1f);
    const float sy = p * y;
    float q = 0x1.d9cce2p-3f;
    q = fmaf(q, t, -0x1.55c5b0p+0f);
    q = fmaf(q, t, 0x1.03c1dcp+2f);
    q = fmaf(q, t, -0x1.3bd3ccp+2f);
    const float cy = fmaf(q, t, 1.0f);
    s = copysignf(swap ? cy : sy, x);
    c = swap ? sy : cy;
}

// sin(2 pi u) and cos(2 pi u) for u in [0, 1) a multiple of 2^-24 (a full
// turn): x = 2u - 1 and, past |x| = 1/2, z = 1 - |x| are exact, so the error
// is that of sincospi_half; (sin, cos)(2 pi u) = -(sin, cos)(pi x).
__device__ __forceinline__ void sincospi_turn(float u, float &s, float &c)
{
    const float x = 2.0f * u - 1.0f;
    const float a = fabsf(x);
    const bool far = a > 0.5f;
    float sz, cz;
    sincospi_half(far ? 1.0f - a : a, sz, cz);
    s = -copysignf(sz, x);
    c = far ? cz : -cz;
}

__device__ __forceinline__ Draw make_draw(uint64_t seed, uint32_t stream, uint64_t path, uint32_t k, uint32_t N,
      uint32_t maxDiff)
{
    const U4 r = philox4x32_10(k, stream, (uint32_t)path, (uint32_t)(path >> 32), (uint32_t)seed,
          (uint32_t)(seed >> 32));
    Draw d;
    d.diff = 2u + __umulhi(r.x, maxDiff - 1u);
    d.left = 1u + __umulhi(r.y, N - d.diff - 1u);
    const int32_t odd = (int32_t)(((r.z >> 8) << 1) | 1u) - (int32_t)(1u << 24);
    const float x = (float)odd * 0x1p-25f;
    sincospi_half(x, d.sh, d.ch);
    return d;
}

// Rotation matrix of the quaternion (cos(theta/2), n sin(theta/2)) from
// c = cos theta, s = sin theta, C = 1 - cos theta: active rotation by theta
// about n, identical to the sandwich q v q* of CurvePerturbUtils.cpp:113-127.
struct Mat3 {
    float m00, m01, m02, m10, m11, m12, m20, m21, m22;
};

__device__ __forceinline__ Mat3 rotation_from_axis(float nx, float ny, float nz, float c, float s, float C)
{
    const float xC = nx * C, yC = ny * C, zC = nz * C;
    const float xs = nx * s, ys = ny * s, zs = nz * s;
    Mat3 R;
    R.m00 = fmaf(nx, xC, c);
    R.m01 = fmaf(nx, yC, -zs);
    R.m02 = fmaf(nx, zC, ys);
    R.m10 = fmaf(ny, xC, zs);
    R.m11 = fmaf(ny, yC, c);
    R.m12 = fmaf(ny, zC, -xs);
    R.m20 = fmaf(nz, xC, -ys);
    R.m21 = fmaf(nz, yC, xs);
    R.m22 = fmaf(nz, zC, c);
    return R;
}

// The same rotation in Rodrigues' form, v c + (n x v) s + n (n . v) C, with
// n = a / |a| folded into the factors (a as summed, not normalised).  Per
// perturbation 7 operations and per segment 19, against 22 and 9 for the
// matrix: the cheaper form where a lane rotates one segment (groups of
// G >= 25 lanes).  Equal to the matrix in exact arithmetic; in FP32 both are
// within a few ulp of it.
struct AxisRotation {
    float ax, ay, az;   // the axis as summed
    float c, sInv, CInv2;  // cos, sin / |a|, (1 - cos) / |a|^2
};

__device__ __forceinline__ AxisRotation axis_rotation(float ax, float ay, float az, float c, float s, float C)
{
    const float inv = __builtin_amdgcn_rsqf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
    return { ax, ay, az, c, s * inv, (C * inv) * inv };
}

__device__ __forceinline__ void rotate_rodrigues(const AxisRotation &R, float &x, float &y, float &z)
{
    const float cx = fmaf(R.ay, z, -(R.az * y));
    const float cy = fmaf(R.az, x, -(R.ax * z));
    const float cz = fmaf(R.ax, y, -(R.ay * x));
    const float k = R.CInv2 * fmaf(R.ax, x, fmaf(R.ay, y, R.az * z));
    const float rx = fmaf(R.c, x, fmaf(R.sInv, cx, k * R.ax));
    const float ry = fmaf(R.c, y, fmaf(R.sInv, cy, k * R.ay));
    const float rz = fmaf(R.c, z, fmaf(R.sInv, cz, k * R.az));
    x = rx;
    y = ry;
    z = rz;
}

// ---------------------------------------------------------------------------
// Segment weight lookup.  PathWeighters.cpp:41-82 with the reference's float
// rounding kept step by step (no contraction), so identical curvatures give
// identical lerped weights: snap/clamp (:45-65), floor index with the last-bin
// fix (:67-74), fraction in curvature units (:79), float lerp (:81-82).  The
// bin index is clamped into the table for curvatures that round past the last
// bin (the reference would read one past the end there).
// ---------------------------------------------------------------------------
struct TableView {
    const float *w;  // LDS or global memory
    int32_t size;
    float minK, maxK, step;
    float invStep;   // RN(1 / step), for the bin index (bin_floor)
};

__device__ __forceinline__ TableView make_table_view(const float *w, int32_t size, float minK, float maxK, float step)
{
    return { w, size, minK, maxK, step, __fdiv_rn(1.0f, step) };
}

// floor(RN(d / step)) for 0 <= d (NaN allowed), the bin index of
// PathWeighters.cpp:67-68, without a division per joint.  q = RN(d * RN(1/step))
// is within (1 + 2^-24)^2 - 1 < 1.51 * 2^-23 of d/step relative, and RN(d/step)
// within 2^-24, so |q - RN(d/step)| < q * 2^-22 for q in the normal range
// (below it both are far under 1/2 and floor to 0).  Unless q is that close to
// an integer (or d is NaN) no integer lies between the two, and floor(q) is
// floor(RN(d/step)); d = 0 gives 0 either way.  The rare remaining lanes (about
// 2e-5 of the curvatures) take the IEEE division, behind a wave-uniform branch
// so that the wave skips it unless one of its lanes needs it (a per-lane
// condition is if-converted: the division's dozen instructions would run for
// every joint).  tests/test_bin_floor.py checks the bound and the floors for
// every float d of the workloads' table ranges.
__device__ __forceinline__ float bin_floor(const TableView &t, float d)
{
    float q = __fmul_rn(d, t.invStep);
    const float k = rintf(q);
    const bool slow = !((d == 0.0f) | (fabsf(q - k) > q * 0x1p-22f));
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(slow) != 0, 0)) {
        const float r = __fdiv_rn(d, t.step);
        q = slow ? r : q;
    }
    return floorf(q);
}

// Curvature moved onto [minK, maxK] as PathWeighters.cpp:45-65 does it (a
// first clamp to minK cannot trigger the second), without branches; *moved
// is set when either bound applied and *counted when the reference would print
// its out-of-range message (more than 1e-3 outside).
__device__ __forceinline__ float clamp_curvature(const TableView &t, float k, bool &moved, bool &counted)
{
    const bool lo = k < t.minK;
    const bool hi = k > t.maxK;
    const bool farLo = lo & !(fabsf(k - t.minK) < 1e-3f);
    const bool farHi = hi & !(fabsf(t.maxK - k) < 1e-3f);
    counted = farLo | farHi;
    moved = lo | hi;
    return lo ? t.minK : (hi ? t.maxK : k);
}

// The interpolation terms of PathWeighters.cpp:67-82 for a curvature on the
// table range: left and right bin, fraction in curvature units, lerped weight.
struct LookupTerms {
    int32_t left, right;
    float frac;   // fraction in curvature units (PathWeighters.cpp:79)
    float value;  // lerped weight
    bool clamped;
};

__device__ __forceinline__ void lerp_terms(const TableView &t, float curvature, LookupTerms &r)
{
    const float distanceFromMin = __fsub_rn(curvature, t.minK);
    int32_t leftIdx = (int32_t)bin_floor(t, distanceFromMin);
    leftIdx = min(max(leftIdx, 0), t.size - 1);
    r.left = leftIdx;
    r.right = (leftIdx == t.size - 1) ? leftIdx : leftIdx + 1;
    r.frac = __fsub_rn(distanceFromMin, __fmul_rn((float)leftIdx, t.step));
    const float a = __fmul_rn(t.w[r.left], __fsub_rn(1.0f, r.frac));
    const float b = __fmul_rn(t.w[r.right], r.frac);
    r.value = __fadd_rn(a, b);
}

__device__ __forceinline__ float lookup_weight(const TableView &t, float curvature, uint32_t &clamped)
{
    bool moved, counted;
    curvature = clamp_curvature(t, curvature, moved, counted);
    clamped += counted ? 1u : 0u;
    LookupTerms r;
    lerp_terms(t, curvature, r);
    return r.value;
}

// The same lookup, returning the interpolation terms the adjoint needs.  value
// is bit-identical to lookup_weight's result for the same curvature; clamped is
// set when the curvature was moved onto the table range (the weight is then
// flat in the curvature).
__device__ __forceinline__ LookupTerms lookup_terms(const TableView &t, float curvature)
{
    LookupTerms r;
    bool counted;
    curvature = clamp_curvature(t, curvature, r.clamped, counted);
    lerp_terms(t, curvature, r);
    return r;
}

// ---------------------------------------------------------------------------
// Compensated float arithmetic.
// ---------------------------------------------------------------------------
struct FF {  // unevaluated sum hi + lo
    float hi, lo;
};

__device__ __forceinline__ FF two_sum(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return { s, e };
}

__device__ __forceinline__ FF ff_add(FF a, FF b)
{
    FF s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    const float hi = s.hi + s.lo;
    return { hi, s.lo - (hi - s.hi) };
}

__device__ __forceinline__ FF ff_add_f(FF a, float b)
{
    FF s = two_sum(a.hi, b);
    s.lo += a.lo;
    const float hi = s.hi + s.lo;
    return { hi, s.lo - (hi - s.hi) };
}

// ---------------------------------------------------------------------------
// Log-domain accumulator in base 2: represents sum_p 2^{v_p} as 2^m * s and
// sum_p 2^{2 v_p} as 2^{2m} * s2, with s and s2 compensated.  This replaces the
// reference's CombinedWeightValues_C (CombinedWeightUtils.h:28-158), which keeps
// a double running total per 1e6 paths with a decimal offset; both carry the
// same quantity without overflow.
// ---------------------------------------------------------------------------
struct LseAcc {
    float m;   // running max of v (-inf when empty)
    FF s;      // sum 2^{v-m}
    FF s2;     // sum 2^{2(v-m)}
    uint32_t nValid, nInvalid, nClamped, pad;
};

__device__ __forceinline__ void lse_init(LseAcc &a)
{
    a.m = -INFINITY;
    a.s = { 0.0f, 0.0f };
    a.s2 = { 0.0f, 0.0f };
    a.nValid = a.nInvalid = a.nClamped = a.pad = 0;
}

// Adds one path with log2 weight v = vhi + vlo.
__device__ __forceinline__ void lse_add(LseAcc &a, float vhi, float vlo)
{
    if (vhi == -INFINITY) return;  // zero weight: counted as valid by the caller
    if (vhi > a.m) {
        const float sc = (a.m == -INFINITY) ? 0.0f : exp2f(a.m - vhi);
        const float sc2 = sc * sc;
        a.s.hi *= sc;
        a.s.lo *= sc;
        a.s2.hi *= sc2;
        a.s2.lo *= sc2;
        a.m = vhi;
    }
    const float d = (vhi - a.m) + vlo;
    const float e = exp2f(d);
    a.s = ff_add_f(a.s, e);
    a.s2 = ff_add_f(a.s2, e * e);
}

__device__ __forceinline__ void lse_merge(LseAcc &a, const LseAcc &b)
{
    a.nValid += b.nValid;
    a.nInvalid += b.nInvalid;
    a.nClamped += b.nClamped;
    if (b.m == -INFINITY) return;
    if (a.m == -INFINITY) {
        a.m = b.m;
        a.s = b.s;
        a.s2 = b.s2;
        return;
    }
    const float M = fmaxf(a.m, b.m);
    const float fa = exp2f(a.m - M), fb = exp2f(b.m - M);
    FF sa { a.s.hi * fa, a.s.lo * fa }, sb { b.s.hi * fb, b.s.lo * fb };
    FF qa { a.s2.hi * fa * fa, a.s2.lo * fa * fa }, qb { b.s2.hi * fb * fb, b.s2.lo * fb * fb };
    a.m = M;
    a.s = ff_add(sa, sb);
    a.s2 = ff_add(qa, qb);
}

// Per-block partial written to global memory; combined in FP64 by
// finalize_kernel.  m is also the block's largest path log2 weight.
struct BlockPartial {
    float m;
    float s_hi, s_lo;
    float s2_hi, s2_lo;
    uint32_t nValid, nInvalid, nClamped;
};

// ---------------------------------------------------------------------------
// Cross-lane sums inside aligned groups of G lanes.  G <= 16 uses DPP row
// patterns (xor 1, xor 2 inside a quad, then half-row and row mirrors, after
// which every lane of the group holds the group total); larger groups finish
// with a permlane swap (32) and a shuffle (64).
// ---------------------------------------------------------------------------
template<int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template<int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

constexpr int DPP_QUAD_XOR1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

// The 32-lane step combines row pairs with gfx950's v_permlane16_swap: called
// with the same register twice, the first result holds the even row of each
// pair and the second the odd one, so their sum is the pair total in every
// lane, in the same order everywhere (every lane of a row already holds the row
// total).  The 64-lane step uses a bpermute shuffle: v_permlane32_swap measured
// slower there.
struct Halves {
    uint32_t even, odd;
};

__device__ __forceinline__ Halves row_pair(uint32_t v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return { r[0], r[1] };
}

template<int G>
__device__ __forceinline__ float group_sum(float v)
{
    if constexpr (G >= 2) v += dpp_f<DPP_QUAD_XOR1>(v);
    if constexpr (G >= 4) v += dpp_f<DPP_QUAD_XOR2>(v);
    if constexpr (G >= 8) v += dpp_f<DPP_ROW_HALF_MIRROR>(v);
    if constexpr (G >= 16) v += dpp_f<DPP_ROW_MIRROR>(v);
    if constexpr (G >= 32) {
        const Halves h = row_pair(__float_as_uint(v));
        v = __uint_as_float(h.even) + __uint_as_float(h.odd);
    }
    if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Lane i ^ 16 within each 32-lane half (ds_swizzle in bit mode: and_mask 0x1F,
// or_mask 0, xor_mask 0x10).  It goes through the LDS crossbar (no LDS
// storage), so it issues no VALU instruction.
__device__ __forceinline__ float swap16(float v)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}

// group_sum of NV independent values, step by step across the values so that
// the DPP steps of different values fill each other's hazard slots.  Every
// lane gets the same totals as group_sum<G> would give (the same additions in
// the same order; the 32-lane step adds its partner row's total, a commutative
// addition of the same two values).
template<int G, int NV>
__device__ __forceinline__ void group_sum_n(float (&v)[NV])
{
    if constexpr (G >= 2) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += dpp_f<DPP_QUAD_XOR1>(v[k]);
    }
    if constexpr (G >= 4) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += dpp_f<DPP_QUAD_XOR2>(v[k]);
    }
    if constexpr (G >= 8) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += dpp_f<DPP_ROW_HALF_MIRROR>(v[k]);
    }
    if constexpr (G >= 16) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += dpp_f<DPP_ROW_MIRROR>(v[k]);
    }
    if constexpr (G >= 32) {
        float o[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) o[k] = swap16(v[k]);
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += o[k];
    }
    if constexpr (G >= 64) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], 32, 64);
    }
}

template<int G>
__device__ __forceinline__ uint32_t group_or(uint32_t v)
{
    if constexpr (G >= 2) v |= dpp_u<DPP_QUAD_XOR1>(v);
    if constexpr (G >= 4) v |= dpp_u<DPP_QUAD_XOR2>(v);
    if constexpr (G >= 8) v |= dpp_u<DPP_ROW_HALF_MIRROR>(v);
    if constexpr (G >= 16) v |= dpp_u<DPP_ROW_MIRROR>(v);
    if constexpr (G >= 32) {
        const Halves h = row_pair(v);
        v = h.even | h.odd;
    }
    if constexpr (G >= 64) v |= (uint32_t)__shfl_xor((int)v, 32, 64);
    return v;
}

template<int G>
__device__ __forceinline__ FF group_sum_ff(FF v)
{
    // Exchanges (hi, lo) pairs with the same partner pattern as group_sum.
    if constexpr (G >= 2) v = ff_add(v, FF { dpp_f<DPP_QUAD_XOR1>(v.hi), dpp_f<DPP_QUAD_XOR1>(v.lo) });
    if constexpr (G >= 4) v = ff_add(v, FF { dpp_f<DPP_QUAD_XOR2>(v.hi), dpp_f<DPP_QUAD_XOR2>(v.lo) });
    if constexpr (G >= 8)
        v = ff_add(v, FF { dpp_f<DPP_ROW_HALF_MIRROR>(v.hi), dpp_f<DPP_ROW_HALF_MIRROR>(v.lo) });
    if constexpr (G >= 16) v = ff_add(v, FF { dpp_f<DPP_ROW_MIRROR>(v.hi), dpp_f<DPP_ROW_MIRROR>(v.lo) });
    if constexpr (G >= 32) {
        const Halves h = row_pair(__float_as_uint(v.hi)), l = row_pair(__float_as_uint(v.lo));
        v = ff_add(FF { __uint_as_float(h.even), __uint_as_float(l.even) },
              FF { __uint_as_float(h.odd), __uint_as_float(l.odd) });
    }
    if constexpr (G >= 64) v = ff_add(v, FF { __shfl_xor(v.hi, 32, 64), __shfl_xor(v.lo, 32, 64) });
    return v;
}

// ---------------------------------------------------------------------------
// LDS helpers shared by the path kernels.
// ---------------------------------------------------------------------------
// Orders one wave's LDS writes before its later LDS reads (lanes of one group
// exchange data through LDS; a group never spans two waves).
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ DrawRec load_rec(const uint32_t *p)
{
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return { v.x, __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w) };
}

// LDS accesses by byte address.  The dynamic-LDS block starts right after the
// static LDS, at __builtin_amdgcn_groupstaticsize(); addresses built from it are
// plain integers, so a slot's row offset stays an instruction immediate.
typedef __attribute__((address_space(3))) float lds_f32;

__device__ __forceinline__ float lds_ld(uint32_t byteAddr) { return *(lds_f32 *)(size_t)byteAddr; }
__device__ __forceinline__ void lds_st(uint32_t byteAddr, float v) { *(lds_f32 *)(size_t)byteAddr = v; }

// 1 / |a| of a segment vector as the forward pass forms it.
__device__ __forceinline__ float inv_norm(float a0, float a1, float a2)
{
    return __builtin_amdgcn_rsqf(fmaf(a0, a0, fmaf(a1, a1, __fmul_rn(a2, a2))));
}

// Unit tangent of a segment vector as the forward pass forms it.
__device__ __forceinline__ float unit_tangent(float a0, float a1, float a2, float &t0, float &t1, float &t2)
{
    const float ia = inv_norm(a0, a1, a2);
    t0 = __fmul_rn(a0, ia);
    t1 = __fmul_rn(a1, ia);
    t2 = __fmul_rn(a2, ia);
    return ia;
}

// The joint between segment a (unit tangent ta, rounded) and segment b (raw,
// ib = 1/|b|): e = (t_b - t_a) N/L and the curvature kappa = |e|
// (CurvePerturbUtils.cpp:219-243).  The difference is formed as
// fma(b, ib, -ta), one rounding, then scaled: written out so that every kernel
// forms the same bits whatever the compiler would contract.  kappa is the
// hardware square root (v_sqrt_f32, one instruction): the IEEE-rounded sqrtf
// costs about 17 more per joint (denormal scaling and two correction steps).
// Measured on gfx950 over every non-negative float: on normal inputs 15% of the
// results differ from the IEEE ones, each by one unit in the last place; inputs
// below 2^-126 (kappa < 1.1e-19, far inside the table's first bin) give 0.
struct JointDiff {
    float e0, e1, e2, kappa;
};

__device__ __forceinline__ JointDiff joint_diff(float ta0, float ta1, float ta2, float b0, float b1, float b2,
      float ib, float invDs)
{
    JointDiff d;
    d.e0 = __fmul_rn(fmaf(b0, ib, -ta0), invDs);
    d.e1 = __fmul_rn(fmaf(b1, ib, -ta1), invDs);
    d.e2 = __fmul_rn(fmaf(b2, ib, -ta2), invDs);
    d.kappa = __builtin_amdgcn_sqrtf(fmaf(d.e0, d.e0, fmaf(d.e1, d.e1, __fmul_rn(d.e2, d.e2))));
    return d;
}

}  // namespace twisty_hip
