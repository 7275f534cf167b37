// path_kernels.hpp -- the per-perturbation step shared by the path kernels
// (plain forward and batched).  Included by .hip files only.
#pragma once

#include "device_common.hpp"
#include "forward.hpp"

namespace twisty_hip {

// One perturbation applied by lane j of its group: read this lane's segments of
// the sub-chain [left, right), sum them across the group into the axis
// p_right - p_left, build the rotation and write the rotated segments back.
// Window slots past this lane's count are redirected to the lane's zeroed
// scratch row: they read zeros, add nothing to the axis and write zeros back,
// so every load and store is unconditional with a compile-time offset.
// scrB[w] is the byte address that lands on the scratch row after slot w's
// offset is added.  With SPLIT the last slot, needed only by a lane holding
// WMAX segments of the sub-chain (diff = 25 for G = 4: one path in 24), sits
// behind a wave-uniform branch taken when some lane of the wave needs it; it
// is loaded and added to the axis last (the same summation order as an
// unconditional slot) and reloaded for the rotation, so no window register
// stays live across the branch.  That pays at 12 waves per CU (lean forward,
// C2: about 2% faster) and costs 1-4% at 8 (table-in-LDS forward, batched),
// so only the lean kernels use it.
template<int G, int BLOCK, bool SPLIT = false>
__device__ __forceinline__ void perturb_one(uint32_t laneB, uint32_t j,
      const uint32_t (&scrB)[(25 + G - 1) / G], const DrawRec &rec)
{
    constexpr int WMAX = (25 + G - 1) / G;       // sub-chain slots per lane (diff <= 25)
    constexpr int WMAIN = (SPLIT && WMAX > 1) ? WMAX - 1 : WMAX;  // slots every perturbation runs
    constexpr uint32_t SLOTB = 12u * BLOCK;      // bytes per slot row
    constexpr uint32_t COMPB = 4u * BLOCK;       // bytes between components
    const uint32_t diff = rec.packed & 0xFFFFu;
    const uint32_t left = rec.packed >> 16;
    // This lane owns segments s G + j.  Its first slot at or past left is
    // floor((left - 1 - j) / G) + 1 and the number of its slots in
    // [left, left + diff) is floor((left + diff - 1 - j) / G) - floor((left -
    // 1 - j) / G) (arithmetic shifts: left - 1 - j >= -G).
    constexpr int LOG2G = __builtin_ctz((unsigned)G);
    const int32_t a = (int32_t)left - (int32_t)(j + 1u);
    const int32_t sa = a >> LOG2G;
    const uint32_t cnt = (uint32_t)(((a + (int32_t)diff) >> LOG2G) - sa);
    const uint32_t baseB = (uint32_t)(sa + 1) * SLOTB + laneB;
    uint32_t addr[WMAIN];
    float dx[WMAIN], dy[WMAIN], dz[WMAIN];
#pragma unroll
    for (int w = 0; w < WMAIN; ++w) {
        addr[w] = ((uint32_t)w < cnt) ? baseB : scrB[w];
        const uint32_t a = addr[w] + w * SLOTB;
        dx[w] = lds_ld(a);
        dy[w] = lds_ld(a + COMPB);
        dz[w] = lds_ld(a + 2 * COMPB);
    }
    float ax = dx[0], ay = dy[0], az = dz[0];
#pragma unroll
    for (int w = 1; w < WMAIN; ++w) {
        ax += dx[w];
        ay += dy[w];
        az += dz[w];
    }
    // wave-uniform: some lane of the wave holds WMAX segments of its sub-chain
    const bool last = WMAIN < WMAX && __builtin_amdgcn_ballot_w64(cnt > (uint32_t)WMAIN) != 0;
    const uint32_t lastA = ((uint32_t)WMAIN < cnt ? baseB : scrB[WMAX - 1]) + (uint32_t)WMAIN * SLOTB;
    if (last) {
        ax += lds_ld(lastA);
        ay += lds_ld(lastA + COMPB);
        az += lds_ld(lastA + 2 * COMPB);
    }
    if constexpr (G > 1) {
        ax = group_sum<G>(ax);
        ay = group_sum<G>(ay);
        az = group_sum<G>(az);
    }
    // The axis is the chord of >= 2 segments; it vanishes only for a sub-chain
    // that closes exactly, and then the path turns NaN and is counted invalid.
    if constexpr (WMAX == 1) {
        // one slot per lane: Rodrigues' form (the one perturb_multi uses)
        const AxisRotation R = axis_rotation(ax, ay, az, rec.c, rec.s, rec.C);
        float x = dx[0], y = dy[0], z = dz[0];
        rotate_rodrigues(R, x, y, z);
        lds_st(addr[0], x);
        lds_st(addr[0] + COMPB, y);
        lds_st(addr[0] + 2 * COMPB, z);
        return;
    }
    const float inv = __builtin_amdgcn_rsqf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
    const Mat3 R = rotation_from_axis(ax * inv, ay * inv, az * inv, rec.c, rec.s, rec.C);
#pragma unroll
    for (int w = 0; w < WMAIN; ++w) {
        const float x = dx[w], y = dy[w], z = dz[w];
        const uint32_t a = addr[w] + w * SLOTB;
        lds_st(a, fmaf(R.m00, x, fmaf(R.m01, y, R.m02 * z)));
        lds_st(a + COMPB, fmaf(R.m10, x, fmaf(R.m11, y, R.m12 * z)));
        lds_st(a + 2 * COMPB, fmaf(R.m20, x, fmaf(R.m21, y, R.m22 * z)));
    }
    if (last) {
        const float x = lds_ld(lastA), y = lds_ld(lastA + COMPB), z = lds_ld(lastA + 2 * COMPB);
        lds_st(lastA, fmaf(R.m00, x, fmaf(R.m01, y, R.m02 * z)));
        lds_st(lastA + COMPB, fmaf(R.m10, x, fmaf(R.m11, y, R.m12 * z)));
        lds_st(lastA + 2 * COMPB, fmaf(R.m20, x, fmaf(R.m21, y, R.m22 * z)));
    }
}


// Whether the sub-chains [left, left + diff) of M records are pairwise
// disjoint (uniform across a group: every lane holds the same records).
template<int M>
__device__ __forceinline__ bool disjoint_records(const DrawRec (&r)[M])
{
    bool ok = true;  // bitwise, not short-circuit: no branches
#pragma unroll
    for (int a = 0; a < M; ++a) {
#pragma unroll
        for (int b = a + 1; b < M; ++b) {
            const uint32_t la = r[a].packed >> 16, lb = r[b].packed >> 16;
            const uint32_t ra = la + (r[a].packed & 0xFFFFu), rb = lb + (r[b].packed & 0xFFFFu);
            ok = ok & ((ra <= lb) | (rb <= la));
        }
    }
    return ok;
}

// M perturbations of one path with pairwise disjoint sub-chains, applied at
// once by a group of G >= 25 lanes (every lane holds at most one segment of a
// sub-chain, so the window is one slot).  Disjoint perturbations read and
// write disjoint segments and their draws do not depend on the state, so the
// result is bit for bit that of perturb_one applied to them in order (the same
// operations on the same values); the M independent chains of LDS loads, group
// sums and rotations interleave and hide each other's latency.  A lane outside
// a sub-chain reads and writes its zeroed scratch row at scrB0, as in
// perturb_one: every redirected slot of the M loads reads zero (the loads all
// precede the stores) and stores R * 0 = 0.
template<int G, int BLOCK, int M>
__device__ __forceinline__ void perturb_multi(uint32_t laneB, uint32_t j, uint32_t scrB0, const DrawRec (&rec)[M])
{
    static_assert(G >= 25, "one window slot per lane needs G >= 25");
    constexpr uint32_t SLOTB = 12u * BLOCK;
    constexpr uint32_t COMPB = 4u * BLOCK;
    uint32_t addr[M];
    float dx[M], dy[M], dz[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const uint32_t diff = rec[m].packed & 0xFFFFu;
        const uint32_t left = rec[m].packed >> 16;
        const uint32_t i0 = left + ((j - left) & (uint32_t)(G - 1));
        addr[m] = (i0 < left + diff) ? (i0 / G) * SLOTB + laneB : scrB0;
        dx[m] = lds_ld(addr[m]);
        dy[m] = lds_ld(addr[m] + COMPB);
        dz[m] = lds_ld(addr[m] + 2 * COMPB);
    }
    float a[3 * M];  // the M axes, reduced together
#pragma unroll
    for (int m = 0; m < M; ++m) {
        a[3 * m] = dx[m];
        a[3 * m + 1] = dy[m];
        a[3 * m + 2] = dz[m];
    }
    group_sum_n<G, 3 * M>(a);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const AxisRotation R = axis_rotation(a[3 * m], a[3 * m + 1], a[3 * m + 2], rec[m].c, rec[m].s, rec[m].C);
        float x = dx[m], y = dy[m], z = dz[m];
        rotate_rodrigues(R, x, y, z);
        lds_st(addr[m], x);
        lds_st(addr[m] + COMPB, y);
        lds_st(addr[m] + 2 * COMPB, z);
    }
}

// Bits of the draw records set by mark_disjoint_chunks (diff <= 25 uses bits
// 0-4): on a chunk's first record when the chunk's M sub-chains are pairwise
// disjoint, and on the first record of each half chunk when that half's are.
constexpr uint32_t kChunkDisjoint = 0x8000u;
constexpr uint32_t kHalfDisjoint = 0x4000u;

// Whether the sub-chains of the packed records pk[lo..hi) are pairwise
// disjoint (bitwise, no branches).
template<int LO, int HI, int N>
__device__ __forceinline__ bool packed_disjoint(const uint32_t (&pk)[N])
{
    bool ok = true;
#pragma unroll
    for (int a = LO; a < HI; ++a) {
#pragma unroll
        for (int b = a + 1; b < HI; ++b) {
            const uint32_t la = pk[a] >> 16, lb = pk[b] >> 16;
            const uint32_t ra = la + (pk[a] & 0xFFFFu), rb = lb + (pk[b] & 0xFFFFu);
            ok = ok & ((ra <= lb) | (rb <= la));
        }
    }
    return ok;
}

// Once per batch, before perturb_batch_multi<..., FLAGGED = true>: lane j of a
// group with j mod M == 0 checks the chunk of records j .. j + M - 1 and marks
// the chunk (all M drawn: j + M <= kcount) and each half of it (its records
// drawn) that is pairwise disjoint, so the group reads bits instead of each
// lane repeating the comparisons.  The records sit in the group's slice of
// drawbuf; the caller syncs the wave before and after.
template<int G, int M>
__device__ __forceinline__ void mark_disjoint_chunks(uint32_t *drawbuf, uint32_t groupBase, uint32_t j,
      uint32_t kcount)
{
    constexpr int H = M / 2;
    if ((j % (uint32_t)M) == 0 && j + (uint32_t)H <= kcount) {
        uint32_t pk[M];
#pragma unroll
        for (int t = 0; t < M; ++t) pk[t] = drawbuf[(groupBase + j + (uint32_t)t) * 4];  // past kcount: unused
        const bool full = j + (uint32_t)M <= kcount;
        const bool half0 = packed_disjoint<0, H>(pk);
        const bool half1 = full && packed_disjoint<H, M>(pk);
        const bool all = full && packed_disjoint<0, M>(pk);
        drawbuf[(groupBase + j) * 4] = pk[0] | (all ? kChunkDisjoint : 0u) | (half0 ? kHalfDisjoint : 0u);
        if (half1) drawbuf[(groupBase + j + (uint32_t)H) * 4] = pk[H] | kHalfDisjoint;
    }
}

// Perturbations kk0 .. kk0 + kcount - 1 of a batch for G >= 25, M at a time:
// a chunk of M whose sub-chains are pairwise disjoint in every group of the
// wave runs as one perturb_multi, any other chunk one perturbation at a time
// or, FLAGGED (the chunks marked by mark_disjoint_chunks), one half chunk at a
// time where a half is disjoint in every group of the wave.  rec(t) returns
// the record of perturbation t of the batch (t < G).
template<int G, int BLOCK, bool SPLIT, int M, bool FLAGGED, class RecOf>
__device__ __forceinline__ void perturb_batch_multi(uint32_t laneB, uint32_t j,
      const uint32_t (&scrB)[(25 + G - 1) / G], uint32_t kcount, RecOf rec)
{
    constexpr int H = M / 2;
    for (uint32_t kk = 0; kk < kcount; kk += M) {
        DrawRec r[M];
#pragma unroll
        for (int t = 0; t < M; ++t) r[t] = rec(min(kk + (uint32_t)t, (uint32_t)G - 1u));
        bool ok;
        uint32_t halves = 0;  // FLAGGED: bit h = half h marked disjoint
        if constexpr (FLAGGED) {
            ok = (r[0].packed & kChunkDisjoint) != 0u;
            halves = ((r[0].packed & kHalfDisjoint) ? 1u : 0u) | ((r[H].packed & kHalfDisjoint) ? 2u : 0u);
            r[0].packed &= ~(kChunkDisjoint | kHalfDisjoint);
            r[H].packed &= ~kHalfDisjoint;
        } else {
            ok = kcount - kk >= (uint32_t)M && disjoint_records<M>(r);
        }
        if (__builtin_amdgcn_ballot_w64(!ok) == 0) {
            perturb_multi<G, BLOCK, M>(laneB, j, scrB[0], r);
        } else if constexpr (FLAGGED && H >= 2) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (kk + (uint32_t)(h * H) >= kcount) break;
                if (__builtin_amdgcn_ballot_w64(((halves >> h) & 1u) == 0u) == 0) {
                    DrawRec rh[H];
#pragma unroll
                    for (int t = 0; t < H; ++t) rh[t] = r[h * H + t];
                    perturb_multi<G, BLOCK, H>(laneB, j, scrB[0], rh);
                } else {
#pragma unroll
                    for (int t = 0; t < H; ++t)
                        if (kk + (uint32_t)(h * H + t) < kcount)
                            perturb_one<G, BLOCK, SPLIT>(laneB, j, scrB, r[h * H + t]);
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; ++t)
                if (kk + (uint32_t)t < kcount) perturb_one<G, BLOCK, SPLIT>(laneB, j, scrB, r[t]);
        }
    }
}

// Perturbations applied together by perturb_batch_multi (groups of G >= 25).
constexpr int kMultiPerturb = 4;

// Draw record of the lane whose wave-lane byte index is srcB (ds_bpermute: a
// cross-lane read through the LDS crossbar that uses no LDS storage).
__device__ __forceinline__ DrawRec bperm_rec(int srcB, const DrawRec &r)
{
    return { (uint32_t)__builtin_amdgcn_ds_bpermute(srcB, (int)r.packed),
        __int_as_float(__builtin_amdgcn_ds_bpermute(srcB, __float_as_int(r.c))),
        __int_as_float(__builtin_amdgcn_ds_bpermute(srcB, __float_as_int(r.s))),
        __int_as_float(__builtin_amdgcn_ds_bpermute(srcB, __float_as_int(r.C))) };
}

// K perturbations of one path, applied in order by its group of G lanes.
// For G > 1 lane j draws perturbation kb + j of each batch of G and the
// records are exchanged through the group's slice of drawbuf (4 words per
// lane), or with ds_bpermute when LEAN (no drawbuf: the LDS it would take goes
// to more resident paths).
template<int G, int BLOCK, bool LEAN = false>
__device__ __forceinline__ void perturb_path(uint32_t laneB, uint32_t lane, uint32_t j, uint32_t groupBase,
      const uint32_t (&scrB)[(25 + G - 1) / G], uint32_t *drawbuf, uint64_t seed, uint32_t stream, uint64_t path,
      uint32_t K, uint32_t N, uint32_t maxDiff)
{
    if constexpr (LEAN && G > 1 && G <= 8) {
        // Batch kb + G is drawn while batch kb is applied: the draws do not
        // depend on the state, so their Philox rounds can fill the waits of
        // the batch's first perturbation.  The records are applied in the
        // same order; draws past the last perturbation are never read.
        const int firstB = (int)(((lane & 63u) - j) * 4u);  // the group's first lane, bytes
        DrawRec mine = pack_draw(make_draw(seed, stream, path, j, N, maxDiff));
        for (uint32_t kb = 0; kb < K; kb += G) {
            const uint32_t kcount = (K - kb < (uint32_t)G) ? (K - kb) : (uint32_t)G;
            DrawRec recs[G];
#pragma unroll
            for (int kk = 0; kk < G; ++kk) recs[kk] = bperm_rec(firstB + 4 * kk, mine);
            mine = pack_draw(make_draw(seed, stream, path, kb + G + j, N, maxDiff));
#pragma unroll
            for (int kk = 0; kk < G; ++kk) {
                if ((uint32_t)kk >= kcount) break;
                perturb_one<G, BLOCK, true>(laneB, j, scrB, recs[kk]);
            }
        }
        return;
    }
    for (uint32_t kb = 0; kb < K; kb += G) {
        const uint32_t kcount = (K - kb < (uint32_t)G) ? (K - kb) : (uint32_t)G;
        if constexpr (G == 1) {
            perturb_one<G, BLOCK>(laneB, j, scrB, pack_draw(make_draw(seed, stream, path, kb, N, maxDiff)));
        } else if constexpr (LEAN) {
            DrawRec mine { 0u, 0.0f, 0.0f, 0.0f };  // past K: never read (kk < kcount)
            if (kb + j < K) mine = pack_draw(make_draw(seed, stream, path, kb + j, N, maxDiff));
            // byte index of the group's first lane within the wave
            const int firstB = (int)(((lane & 63u) - j) * 4u);
            if constexpr (G >= 25) {  // G <= 8 takes the loop above
                perturb_batch_multi<G, BLOCK, true, kMultiPerturb, false>(laneB, j, scrB, kcount,
                      [&](uint32_t t) { return bperm_rec(firstB + 4 * (int)t, mine); });
            } else {
                for (uint32_t kk = 0; kk < kcount; ++kk)
                    perturb_one<G, BLOCK, true>(laneB, j, scrB, bperm_rec(firstB + 4 * (int)kk, mine));
            }
        } else {
            if (kb + j < K) {
                const DrawRec r = pack_draw(make_draw(seed, stream, path, kb + j, N, maxDiff));
                *reinterpret_cast<uint4 *>(drawbuf + lane * 4) =
                      make_uint4(r.packed, __float_as_uint(r.c), __float_as_uint(r.s), __float_as_uint(r.C));
            }
            wave_sync_lds();
            if constexpr (G <= 8) {
                DrawRec recs[G];
#pragma unroll
                for (int kk = 0; kk < G; ++kk) recs[kk] = load_rec(drawbuf + (groupBase + kk) * 4);
                wave_sync_lds();
#pragma unroll
                for (int kk = 0; kk < G; ++kk) {
                    if ((uint32_t)kk >= kcount) break;
                    perturb_one<G, BLOCK>(laneB, j, scrB, recs[kk]);
                }
            } else if constexpr (G >= 25) {
                mark_disjoint_chunks<G, kMultiPerturb>(drawbuf, groupBase, j, kcount);
                wave_sync_lds();
                perturb_batch_multi<G, BLOCK, false, kMultiPerturb, true>(laneB, j, scrB, kcount,
                      [&](uint32_t t) { return load_rec(drawbuf + (groupBase + t) * 4); });
                wave_sync_lds();
            } else {
                for (uint32_t kk = 0; kk < kcount; ++kk)
                    perturb_one<G, BLOCK>(laneB, j, scrB, load_rec(drawbuf + (groupBase + kk) * 4));
                wave_sync_lds();
            }
        }
    }
}

// Copies this lane's segments of a curve in global memory (component c of
// segment i at src + i*segStride + c*compStride) into its state columns
// (slot s, component c at s*3*BLOCK + c*BLOCK + lane) with direct global ->
// LDS loads: no register round trip, so all of the path's 3*S loads are in
// flight at once instead of one slot per memory latency.  Each wave fills its
// own lanes' columns and waits for its loads before it goes on; the wave's
// earlier reads of these rows have completed (their values were used).
template<int G, int BLOCK>
__device__ __forceinline__ void load_state(float *smem, const float *src, uint64_t segStride, uint64_t compStride,
      uint32_t S, uint32_t N, uint32_t lane, uint32_t j)
{
    constexpr uint32_t SLOT = 3u * BLOCK;
    float *rows = smem + (lane & ~63u);  // the wave's first lane: M0 of the LDS writes
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t i = s * G + j;
        if (i < N) {
            const float *a = src + (uint64_t)i * segStride;
            __builtin_amdgcn_global_load_lds(a, rows + s * SLOT, 4, 0, 0);
            __builtin_amdgcn_global_load_lds(a + compStride, rows + s * SLOT + BLOCK, 4, 0, 0);
            __builtin_amdgcn_global_load_lds(a + 2 * compStride, rows + s * SLOT + 2 * BLOCK, 4, 0, 0);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // every counter, vmcnt included: the LDS writes have landed
    wave_sync_lds();
}

struct PathLog {
    FF lsum;         // sum of log2 segment weights (valid joints)
    uint32_t flags;  // 1 = NaN, 2 = -inf, 4 = +inf segment log
};

// log2 of the weight of joint i (segments i and i+1) of the state in LDS:
// tangents (CurvePerturbUtils.cpp:147-154), curvature (:219-243), lookup
// (PathWeighters.cpp:41-82); segment i sits in lane i mod G of the group.
// One joint's weighting with its by-products: the curvature and the unit
// tangent of its first segment (the exit-direction bin uses t_{N-2}).
struct JointEval {
    float l;            // log2 of the segment weight
    float kappa;
    float t0, t1, t2;   // unit tangent of segment i
};

template<int G, int BLOCK>
__device__ __forceinline__ JointEval joint_eval(const float *smem, uint32_t lane, uint32_t groupBase, uint32_t i,
      float invDs, const TableView &tv, uint32_t &clamped)
{
    constexpr uint32_t SLOT = 3u * BLOCK;
    const uint32_t ai = (i / G) * SLOT + lane;
    JointEval r;
    unit_tangent(smem[ai], smem[ai + BLOCK], smem[ai + 2 * BLOCK], r.t0, r.t1, r.t2);
    const uint32_t i1 = i + 1;
    const uint32_t bi = (i1 / G) * SLOT + groupBase + (i1 & (uint32_t)(G - 1));
    const float b0 = smem[bi], b1 = smem[bi + BLOCK], b2 = smem[bi + 2 * BLOCK];
    r.kappa = joint_diff(r.t0, r.t1, r.t2, b0, b1, b2, inv_norm(b0, b1, b2), invDs).kappa;
    r.l = log2f(lookup_weight(tv, r.kappa, clamped));
    return r;
}

// Per-joint by-products are not wanted.
struct NoJointOutput {
    __device__ __forceinline__ void operator()(uint32_t, const JointEval &) const {}
};

// Adds a joint's log2 weight l to the path's sum, or records it in the flags
// when it is not finite (PathWeighters.cpp:88-96); a joint with use = false
// changes nothing.  The kind of a non-finite log (rare: a zero or non-finite
// weight) is worked out behind a wave-uniform branch.
__device__ __forceinline__ void add_joint_log(FF &lsum, uint32_t &flags, float l, bool use)
{
    const bool finite = __builtin_isfinite(l);
    const FF t = ff_add_f(lsum, l);
    const bool take = use & finite;
    lsum.hi = take ? t.hi : lsum.hi;
    lsum.lo = take ? t.lo : lsum.lo;
    const bool odd = use & !finite;
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(odd) != 0, 0)) {
        const uint32_t f = !(l == l) ? 1u : (l < 0.0f ? 2u : 4u);
        flags |= odd ? f : 0u;
    }
}

// Joints a lane weighs per pass of weigh_state.  One is fastest as measured
// (weighting of 2^24 paths of 1024 segments at 32 lanes per path: 118 ms at
// one joint per pass, 142 ms at two, 127 ms at four; 64 segments at 4 lanes
// also favour one).  At one per pass each joint waits for its own LDS and
// table reads, but the loop is not latency-bound: software-pipelined by hand
// (joint i + G's reads, curvature and bin begun before joint i's lerp and
// log2; the same bits) it took 101.7 ms against 93.8 for that weighting.
constexpr int kJointsPerPass = 1;

// Weighting of the state in LDS (G > 1): lane j takes joints i = j mod G,
// kJointsPerPass per pass (i, i + G, ...); the sum keeps the joint order, so
// every kJointsPerPass gives the same bits.  out(i, joint) receives every
// joint's by-products (per-joint outputs of the forward pass).  Every lane of
// the group returns the group's totals.
template<int G, int BLOCK, class Out = NoJointOutput>
__device__ __forceinline__ PathLog weigh_state(const float *smem, uint32_t lane, uint32_t j, uint32_t groupBase,
      uint32_t N, float invDs, const TableView &tv, uint32_t &clamped, Out out = Out())
{
    static_assert(G > 1, "weigh_state handles groups of two or more lanes");
    constexpr int U = kJointsPerPass;
    FF lsum { 0.0f, 0.0f };
    uint32_t flags = 0;
    for (uint32_t i = j; i + 1 < N; i += U * G) {
        // A slot past the last joint recomputes joint i and its result is
        // discarded.
        JointEval e[U];
        bool use[U];
        uint32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t iu = i + (uint32_t)u * G;
            use[u] = iu + 1 < N;
            c[u] = 0;
            e[u] = joint_eval<G, BLOCK>(smem, lane, groupBase, use[u] ? iu : i, invDs, tv, c[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            add_joint_log(lsum, flags, e[u].l, use[u]);
            clamped += use[u] ? c[u] : 0u;
            if (use[u]) out(i + (uint32_t)u * G, e[u]);
        }
    }
    return { group_sum_ff<G>(lsum), group_or<G>(flags) };
}

// FP64 merge of per-block partials parts[idx[k]] (idx NULL: parts[k]),
// k < n, by one block of 256 threads; fixed order, so deterministic.
__device__ __forceinline__ void finalize_partials(const BlockPartial *parts, const uint32_t *idx, uint32_t n,
      double absorbLog2PerPath, FinalOut *out)
{
    __shared__ double shM[256], shS[256], shS2[256];
    __shared__ unsigned long long shV[256], shI[256], shC[256];
    double M = -INFINITY, S = 0.0, S2 = 0.0;
    unsigned long long nv = 0, ni = 0, nc = 0;
    for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) {
        const BlockPartial q = parts[idx ? idx[b] : b];
        nv += q.nValid;
        ni += q.nInvalid;
        nc += q.nClamped;
        if (q.m == -INFINITY) continue;
        const double m = q.m;
        const double s = (double)q.s_hi + (double)q.s_lo;
        const double s2 = (double)q.s2_hi + (double)q.s2_lo;
        if (m > M) {
            const double f = (M == -INFINITY) ? 0.0 : exp2(M - m);
            S = S * f + s;
            S2 = S2 * f * f + s2;
            M = m;
        } else {
            const double f = exp2(m - M);
            S += s * f;
            S2 += s2 * f * f;
        }
    }
    shM[threadIdx.x] = M;
    shS[threadIdx.x] = S;
    shS2[threadIdx.x] = S2;
    shV[threadIdx.x] = nv;
    shI[threadIdx.x] = ni;
    shC[threadIdx.x] = nc;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint32_t o = threadIdx.x + h;
            const double ma = shM[threadIdx.x], mb = shM[o];
            if (mb != -INFINITY) {
                if (ma == -INFINITY) {
                    shM[threadIdx.x] = mb;
                    shS[threadIdx.x] = shS[o];
                    shS2[threadIdx.x] = shS2[o];
                } else {
                    const double Mx = fmax(ma, mb);
                    const double fa = exp2(ma - Mx), fb = exp2(mb - Mx);
                    shS[threadIdx.x] = shS[threadIdx.x] * fa + shS[o] * fb;
                    shS2[threadIdx.x] = shS2[threadIdx.x] * fa * fa + shS2[o] * fb * fb;
                    shM[threadIdx.x] = Mx;
                }
            }
            shV[threadIdx.x] += shV[o];
            shI[threadIdx.x] += shI[o];
            shC[threadIdx.x] += shC[o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double l2 = 0.30102999566398119521;  // log10(2)
        const double m = shM[0];
        out->maxLog2 = m;
        out->sumScaled = shS[0];
        if (m == -INFINITY) {
            out->log10SumW = -INFINITY;
            out->log10SumW2 = -INFINITY;
            out->maxLog10W = -INFINITY;
        } else {
            out->log10SumW = (m + log2(shS[0]) + absorbLog2PerPath) * l2;
            out->log10SumW2 = (2.0 * m + log2(shS2[0]) + 2.0 * absorbLog2PerPath) * l2;
            out->maxLog10W = (m + absorbLog2PerPath) * l2;
        }
        out->nValid = shV[0];
        out->nInvalid = shI[0];
        out->nClamped = shC[0];
    }
}

// Block-wide merge of the per-lane accumulators into *out; scratch holds one
// LseAcc per wave.  Ends with every thread past a barrier only after thread 0
// has read the scratch (the caller's next barrier).
__device__ __forceinline__ void block_reduce_lse(LseAcc acc, float *scratch, BlockPartial *out)
{
    // A block of a size that is not a multiple of 64 ends in a partial wave:
    // lanes past the block hold nothing and are skipped.
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const bool partner = (threadIdx.x ^ (unsigned)o) < blockDim.x;
        LseAcc b;
        b.m = __shfl_xor(acc.m, o, 64);
        b.s.hi = __shfl_xor(acc.s.hi, o, 64);
        b.s.lo = __shfl_xor(acc.s.lo, o, 64);
        b.s2.hi = __shfl_xor(acc.s2.hi, o, 64);
        b.s2.lo = __shfl_xor(acc.s2.lo, o, 64);
        b.nValid = (uint32_t)__shfl_xor((int)acc.nValid, o, 64);
        b.nInvalid = (uint32_t)__shfl_xor((int)acc.nInvalid, o, 64);
        b.nClamped = (uint32_t)__shfl_xor((int)acc.nClamped, o, 64);
        if (partner) lse_merge(acc, b);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, nWaves = (blockDim.x + 63) >> 6;
    LseAcc *sh = reinterpret_cast<LseAcc *>(scratch);
    if ((threadIdx.x & 63) == 0) sh[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        LseAcc a = sh[0];
        for (int w = 1; w < nWaves; ++w) lse_merge(a, sh[w]);
        BlockPartial p;
        p.m = a.m;
        p.s_hi = a.s.hi;
        p.s_lo = a.s.lo;
        p.s2_hi = a.s2.hi;
        p.s2_lo = a.s2.lo;
        p.nValid = a.nValid;
        p.nInvalid = a.nInvalid;
        p.nClamped = a.nClamped;
        *out = p;
    }
}

}  // namespace twisty_hip
