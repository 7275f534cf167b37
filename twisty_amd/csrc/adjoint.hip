// adjoint.hip -- reverse-mode derivative of the log-estimator (gfx950).
//
// E = log10(sum_p w_p / P) over the valid paths.  With omega_p = w_p / sum w,
//   dE/dx = sum_p omega_p d(ln w_p)/dx / ln 10,
// and ln w_p = sum_joints ln f(kappa_q) + const, f the lerped table weight
// (PathWeighters.cpp:41-85).  The kernel reads what the forward pass left in
// HBM -- each path's final segment vectors and its log2 weight -- plus the
// normaliser (M, S) of the finalize kernel, so omega_p = 2^(log2 w_p - M) / S
// needs no second pass over the paths.
//
// Table entries.  f is linear in T[left], T[right] with coefficients
// (1 - frac, frac), so dE/dT[b] collects omega_p (1 - frac)/f and omega_p frac/f
// over every joint: a scatter-add into a block-private copy of the gradient in
// LDS (LDS float atomics), written out per block and summed in a second kernel.
//
// Initial segments (sweep instances).  dE/dd_i^0 is the adjoint of the K
// perturbations run backwards.  A rotation about its own chord leaves the chord
// unchanged, so each step's axis is recomputed from the post-rotation segments;
// the pre-rotation segments are R^T d, so the state is rewound in place instead
// of being stored.  With lambda = dE/dd after the step,
//   lambda_pre = R^T lambda + dE/da   (segments of the sub-chain),
//   dE/dR = sum lambda d_pre^T,  R = c I + s [n]x + C n n^T,
//   dE/dn_k = s (G_ji - G_ij)_{(i,j,k) cyclic} + C ((G n)_k + (G^T n)_k),
//   dE/da = (I - n n^T) dE/dn / |a|.
// Every path starts from the same initial curve, so the per-path lambda^0 are
// summed over paths (LDS atomics per block).
//
// The weight's derivative in the curvature.  The reference's lerp fraction is
// in curvature units, not bin units (PathWeighters.cpp:79), so inside a bin f
// moves by (T[r] - T[l]) per unit curvature and jumps by about T[r] - T[l] at
// the next bin edge.  The pathwise slope alone would miss the jumps, which
// carry almost all of the change; the adjoint uses the bin-average slope
// (T[r] - T[l]) / step, the derivative of the weight averaged over the
// position inside the bin (in-bin slope plus edge jumps per unit curvature).
// For table entries no such choice arises: f is exactly linear in them.
#include "adjoint.hpp"
#include "path_kernels.hpp"

namespace twisty_hip {

// One perturbation undone by lane j of its group: rewinds this lane's
// segments of the sub-chain and their adjoints (same window and scratch-row
// redirection as the forward's perturb_one; the adjoint rows sit lamB bytes
// after the segment rows).  Every slot of a lane past its count lands on the
// same scratch address; they load the same values and store the same values.
// The last slot (needed only by a lane holding WMAX segments of the sub-chain,
// diff = 25 for G = 8) runs behind a wave-uniform branch taken when some lane
// of the wave needs it, as in the forward; its terms are added after the
// other slots', in the same order as an unconditional slot.
template<int G, int BLOCK>
__device__ __forceinline__ void unperturb_one(uint32_t laneB, uint32_t lamB, uint32_t j,
      const uint32_t (&scrB)[(25 + G - 1) / G], const DrawRec &rec)
{
    constexpr int WALL = (25 + G - 1) / G;
    constexpr int WMAX = WALL > 1 ? WALL - 1 : WALL;  // slots every perturbation runs
    constexpr uint32_t SLOTB = 12u * BLOCK;
    constexpr uint32_t COMPB = 4u * BLOCK;
    const uint32_t diff = rec.packed & 0xFFFFu;
    const uint32_t left = rec.packed >> 16;
    // first slot and slot count of this lane in [left, left + diff), as in
    // perturb_one
    constexpr int LOG2G = __builtin_ctz((unsigned)G);
    const int32_t a = (int32_t)left - (int32_t)(j + 1u);
    const int32_t sa = a >> LOG2G;
    const uint32_t cnt = (uint32_t)(((a + (int32_t)diff) >> LOG2G) - sa);
    const uint32_t baseB = (uint32_t)(sa + 1) * SLOTB + laneB;
    uint32_t addr[WMAX];
    float dx[WMAX], dy[WMAX], dz[WMAX], lx[WMAX], ly[WMAX], lz[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
        addr[w] = ((uint32_t)w < cnt) ? baseB : scrB[w];
        const uint32_t a = addr[w] + w * SLOTB;
        dx[w] = lds_ld(a);
        dy[w] = lds_ld(a + COMPB);
        dz[w] = lds_ld(a + 2 * COMPB);
        lx[w] = lds_ld(a + lamB);
        ly[w] = lds_ld(a + lamB + COMPB);
        lz[w] = lds_ld(a + lamB + 2 * COMPB);
    }
    float ax = dx[0], ay = dy[0], az = dz[0];
#pragma unroll
    for (int w = 1; w < WMAX; ++w) {
        ax += dx[w];
        ay += dy[w];
        az += dz[w];
    }
    // wave-uniform: some lane of the wave holds WALL segments of its sub-chain
    const bool last = WMAX < WALL && __builtin_amdgcn_ballot_w64(cnt > (uint32_t)WMAX) != 0;
    const uint32_t lastA = ((uint32_t)WMAX < cnt ? baseB : scrB[WALL - 1]) + (uint32_t)WMAX * SLOTB;
    float ex = 0.f, ey = 0.f, ez = 0.f, mx = 0.f, my = 0.f, mz = 0.f;  // the last slot's d and lambda
    if (last) {
        ex = lds_ld(lastA);
        ey = lds_ld(lastA + COMPB);
        ez = lds_ld(lastA + 2 * COMPB);
        mx = lds_ld(lastA + lamB);
        my = lds_ld(lastA + lamB + COMPB);
        mz = lds_ld(lastA + lamB + 2 * COMPB);
        ax += ex;
        ay += ey;
        az += ez;
    }
    if constexpr (G > 1) {
        ax = group_sum<G>(ax);
        ay = group_sum<G>(ay);
        az = group_sum<G>(az);
    }
    const float inv = __builtin_amdgcn_rsqf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
    const float nx = ax * inv, ny = ay * inv, nz = az * inv;
    const Mat3 R = rotation_from_axis(nx, ny, nz, rec.c, rec.s, rec.C);
    // With G = sum_w lambda_w d_pre,w^T, dE/dn_k = s (G_ji - G_ij) + C ((G n)_k
    // + (G^T n)_k) is linear in the per-slot terms and s, C, n are the
    // group's, so each lane forms its own share of dE/dn,
    //   -s sum_w lambda_w x p_w + C sum_w (lambda_w (p_w . n) + p_w (lambda_w . n)),
    // and the group sums three values instead of the nine entries of G.
    float cx = 0.f, cy = 0.f, cz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    float px[WMAX], py[WMAX], pz[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
        // d_pre = R^T d_post
        px[w] = fmaf(R.m00, dx[w], fmaf(R.m10, dy[w], R.m20 * dz[w]));
        py[w] = fmaf(R.m01, dx[w], fmaf(R.m11, dy[w], R.m21 * dz[w]));
        pz[w] = fmaf(R.m02, dx[w], fmaf(R.m12, dy[w], R.m22 * dz[w]));
        cx = fmaf(ly[w], pz[w], fmaf(-lz[w], py[w], cx));
        cy = fmaf(lz[w], px[w], fmaf(-lx[w], pz[w], cy));
        cz = fmaf(lx[w], py[w], fmaf(-ly[w], px[w], cz));
        const float pn = fmaf(px[w], nx, fmaf(py[w], ny, pz[w] * nz));
        const float ln = fmaf(lx[w], nx, fmaf(ly[w], ny, lz[w] * nz));
        bx = fmaf(lx[w], pn, fmaf(px[w], ln, bx));
        by = fmaf(ly[w], pn, fmaf(py[w], ln, by));
        bz = fmaf(lz[w], pn, fmaf(pz[w], ln, bz));
    }
    float qx_ = 0.f, qy_ = 0.f, qz_ = 0.f;  // the last slot's d_pre
    if (last) {
        qx_ = fmaf(R.m00, ex, fmaf(R.m10, ey, R.m20 * ez));
        qy_ = fmaf(R.m01, ex, fmaf(R.m11, ey, R.m21 * ez));
        qz_ = fmaf(R.m02, ex, fmaf(R.m12, ey, R.m22 * ez));
        cx = fmaf(my, qz_, fmaf(-mz, qy_, cx));
        cy = fmaf(mz, qx_, fmaf(-mx, qz_, cy));
        cz = fmaf(mx, qy_, fmaf(-my, qx_, cz));
        const float pn = fmaf(qx_, nx, fmaf(qy_, ny, qz_ * nz));
        const float ln = fmaf(mx, nx, fmaf(my, ny, mz * nz));
        bx = fmaf(mx, pn, fmaf(qx_, ln, bx));
        by = fmaf(my, pn, fmaf(qy_, ln, by));
        bz = fmaf(mz, pn, fmaf(qz_, ln, bz));
    }
    float hx = fmaf(-rec.s, cx, rec.C * bx);
    float hy = fmaf(-rec.s, cy, rec.C * by);
    float hz = fmaf(-rec.s, cz, rec.C * bz);
    if constexpr (G > 1) {
        hx = group_sum<G>(hx);
        hy = group_sum<G>(hy);
        hz = group_sum<G>(hz);
    }
    // project out the axis direction and divide by |a|
    const float hn = fmaf(hx, nx, fmaf(hy, ny, hz * nz));
    const float dax = fmaf(-hn, nx, hx) * inv;
    const float day = fmaf(-hn, ny, hy) * inv;
    const float daz = fmaf(-hn, nz, hz) * inv;
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
        const uint32_t a = addr[w] + w * SLOTB;
        // Redirected slots add dE/da to the lane's adjoint scratch row too.  That
        // row only ever meets segment scratch values of zero (R^T 0 = 0), so it
        // adds nothing to dE/dR; it is cleared at every path start.
        const float qx = fmaf(R.m00, lx[w], fmaf(R.m10, ly[w], fmaf(R.m20, lz[w], dax)));
        const float qy = fmaf(R.m01, lx[w], fmaf(R.m11, ly[w], fmaf(R.m21, lz[w], day)));
        const float qz = fmaf(R.m02, lx[w], fmaf(R.m12, ly[w], fmaf(R.m22, lz[w], daz)));
        lds_st(a, px[w]);
        lds_st(a + COMPB, py[w]);
        lds_st(a + 2 * COMPB, pz[w]);
        lds_st(a + lamB, qx);
        lds_st(a + lamB + COMPB, qy);
        lds_st(a + lamB + 2 * COMPB, qz);
    }
    if (last) {
        const float qx = fmaf(R.m00, mx, fmaf(R.m10, my, fmaf(R.m20, mz, dax)));
        const float qy = fmaf(R.m01, mx, fmaf(R.m11, my, fmaf(R.m21, mz, day)));
        const float qz = fmaf(R.m02, mx, fmaf(R.m12, my, fmaf(R.m22, mz, daz)));
        lds_st(lastA, qx_);
        lds_st(lastA + COMPB, qy_);
        lds_st(lastA + 2 * COMPB, qz_);
        lds_st(lastA + lamB, qx);
        lds_st(lastA + lamB + COMPB, qy);
        lds_st(lastA + lamB + 2 * COMPB, qz);
    }
}

// d ln f / d kappa and the joint's curvature-direction vector for the joint
// between unit tangents (ta, tb), plus the table scatter when requested.
struct JointAdj {
    float vx, vy, vz;  // (d ln f / d kappa) * d kappa / d t_b  (= -d/d t_a)
};

// The joint between segment a (unit tangent ta) and segment b (raw, 1/|b| =
// ib), with the forward pass's curvature (joint_diff).
template<bool SCATTER>
__device__ __forceinline__ JointAdj joint_adjoint(const TableView &tv, float invDs, float invStep, float ta0,
      float ta1, float ta2, float b0, float b1, float b2, float ib, float wp, float *gradT)
{
    const JointDiff jd = joint_diff(ta0, ta1, ta2, b0, b1, b2, ib, invDs);
    const float e0 = jd.e0, e1 = jd.e1, e2 = jd.e2, kappa = jd.kappa;
    const LookupTerms lt = lookup_terms(tv, kappa);
    if constexpr (SCATTER) {
        const float cw = wp / lt.value;
        atomicAdd(&gradT[lt.left], cw * (1.0f - lt.frac));
        atomicAdd(&gradT[lt.right], cw * lt.frac);
    }
    const float slope = lt.clamped ? 0.0f : (tv.w[lt.right] - tv.w[lt.left]) * invStep;
    const float g = (kappa > 0.0f) ? slope / lt.value * (invDs / kappa) : 0.0f;
    return { g * e0, g * e1, g * e2 };
}

template<int G, int BLOCK, bool SWEEP>
__global__ void __launch_bounds__(BLOCK) adjoint_kernel(const AdjParams p)
{
    constexpr uint32_t SLOT = 3u * BLOCK;
    constexpr int WMAX = (25 + G - 1) / G;
    extern __shared__ float smem[];
    const uint32_t N = p.N, S = p.S, K = p.K;
    const uint32_t rows = p.scratchRow + 1;
    const uint32_t lamOff = SWEEP ? rows * SLOT : 0u;
    const uint32_t tabOff = rows * SLOT + lamOff;
    const uint32_t gtOff = tabOff + p.tableStride;
    const uint32_t gsOff = gtOff + p.tableStride;
    const uint32_t drawOff = gsOff + (SWEEP ? p.segStride : 0u);
    uint32_t *drawbuf = reinterpret_cast<uint32_t *>(smem + drawOff);
    float *gradT = smem + gtOff;
    float *gradS = smem + gsOff;

    for (int i = threadIdx.x; i < p.tableSize; i += BLOCK) {
        smem[tabOff + i] = p.table[i];
        gradT[i] = 0.0f;
    }
    if constexpr (SWEEP)
        for (uint32_t i = threadIdx.x; i < 3 * N; i += BLOCK) gradS[i] = 0.0f;
    const uint32_t lane = threadIdx.x;
    const uint32_t j = lane & (G - 1);
    const uint32_t groupBase = lane - j;
    constexpr uint32_t groupsPerBlock = BLOCK / G;
    const uint32_t scr = p.scratchRow * SLOT + lane;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        smem[scr + c * BLOCK] = 0.0f;
        if constexpr (SWEEP) smem[lamOff + scr + c * BLOCK] = 0.0f;
    }
    __syncthreads();
    const TableView tv = make_table_view(smem + tabOff, p.tableSize, p.minK, p.maxK, p.step);
    const uint32_t ldsBase = __builtin_amdgcn_groupstaticsize();
    const uint32_t laneB = ldsBase + lane * 4u;
    const uint32_t lamB = lamOff * 4u;
    uint32_t scrB[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) scrB[w] = ldsBase + (scr - (uint32_t)w * SLOT) * 4u;
    const float M = (float)p.fin->maxLog2;
    const bool finiteM = M > -INFINITY && M < INFINITY;
    const uint64_t numVisits = p.pick ? (uint64_t)*p.pickCount : p.numPaths;
    float st[kDirStatsMax];  // direction statistics of this group's paths (group leader)
#pragma unroll
    for (uint32_t k = 0; k < kDirStatsMax; ++k) st[k] = 0.0f;

    for (uint64_t visit = (uint64_t)blockIdx.x * groupsPerBlock + lane / G; visit < numVisits;
          visit += (uint64_t)gridDim.x * groupsPerBlock) {
        const uint64_t local = p.pick ? p.pick[visit].local : visit;
        const float2 lw = p.pathLog2[local];
        // Scaled weight 2^(log2 w - M) in (0, 1]; invalid (NaN) and zero-weight
        // paths, and those below FP32 range relative to the largest, add nothing.
        float wp = (finiteM && lw.x > -INFINITY && lw.x < INFINITY) ? exp2f((lw.x - M) + lw.y) : 0.0f;
        const float w0 = wp;                                   // before the roulette factor
        const float factor = p.pick ? p.pick[visit].factor : 1.0f;
        wp *= factor;
        if (!(wp > 0.0f)) continue;  // uniform across the path's group
        const uint64_t path = p.pathOffset + local;
        load_state<G, BLOCK>(smem, p.finalSeg + local, 3ull * p.numPaths, p.numPaths, S, N, lane, j);
        if constexpr (SWEEP) {
#pragma unroll
            for (int c = 0; c < 3; ++c) smem[lamOff + scr + c * BLOCK] = 0.0f;
        }
        wave_sync_lds();

        // Adjoint of the weighting: lambda_i = wp (I - t t^T) dl/dt_i / |d_i|,
        // dl/dt_i = v_{i-1} - v_i over the joints on either side.  The owner of
        // segment i also scatters joint i's table coefficients.
        for (uint32_t s = 0; s < S; ++s) {
            const uint32_t i = s * G + j;
            if (i >= N) break;
            float t0, t1, t2;
            const float a0 = smem[s * SLOT + lane], a1 = smem[s * SLOT + BLOCK + lane];
            const float a2 = smem[s * SLOT + 2 * BLOCK + lane];
            const float ia = unit_tangent(a0, a1, a2, t0, t1, t2);
            float mx = 0.0f, my = 0.0f, mz = 0.0f;
            if (i >= 1) {
                const uint32_t q = i - 1;
                const uint32_t qi = (q / G) * SLOT + groupBase + (q & (uint32_t)(G - 1));
                float u0, u1, u2;
                unit_tangent(smem[qi], smem[qi + BLOCK], smem[qi + 2 * BLOCK], u0, u1, u2);
                const JointAdj ja = joint_adjoint<false>(tv, p.invDs, p.invStep, u0, u1, u2, a0, a1, a2, ia, wp,
                      gradT);
                mx += ja.vx;
                my += ja.vy;
                mz += ja.vz;
            }
            if (i + 1 < N) {
                const uint32_t q = i + 1;
                const uint32_t qi = (q / G) * SLOT + groupBase + (q & (uint32_t)(G - 1));
                const float b0 = smem[qi], b1 = smem[qi + BLOCK], b2 = smem[qi + 2 * BLOCK];
                const float ib = inv_norm(b0, b1, b2);
                const JointAdj ja = p.tableGrad
                      ? joint_adjoint<true>(tv, p.invDs, p.invStep, t0, t1, t2, b0, b1, b2, ib, wp, gradT)
                      : joint_adjoint<false>(tv, p.invDs, p.invStep, t0, t1, t2, b0, b1, b2, ib, wp, gradT);
                mx -= ja.vx;
                my -= ja.vy;
                mz -= ja.vz;
            }
            if constexpr (SWEEP) {
                const float tm = fmaf(t0, mx, fmaf(t1, my, t2 * mz));
                const float sc = wp * ia;
                smem[lamOff + s * SLOT + lane] = fmaf(-tm, t0, mx) * sc;
                smem[lamOff + s * SLOT + BLOCK + lane] = fmaf(-tm, t1, my) * sc;
                smem[lamOff + s * SLOT + 2 * BLOCK + lane] = fmaf(-tm, t2, mz) * sc;
            }
        }

        if constexpr (SWEEP) {
            wave_sync_lds();
            const uint32_t numBatches = (K + G - 1) / G;
            for (uint32_t bb = numBatches; bb-- > 0;) {
                const uint32_t kb = bb * G;
                const uint32_t kcount = (K - kb < (uint32_t)G) ? (K - kb) : (uint32_t)G;
                if constexpr (G == 1) {
                    unperturb_one<G, BLOCK>(laneB, lamB, j, scrB,
                          pack_draw(make_draw(p.seed, p.stream, path, kb, N, p.maxDiff)));
                } else {
                    if (j < kcount) {
                        const DrawRec r = pack_draw(make_draw(p.seed, p.stream, path, kb + j, N, p.maxDiff));
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
                        for (int kk = G - 1; kk >= 0; --kk) {
                            if ((uint32_t)kk < kcount) unperturb_one<G, BLOCK>(laneB, lamB, j, scrB, recs[kk]);
                        }
                    } else {
                        for (uint32_t kk = kcount; kk-- > 0;)
                            unperturb_one<G, BLOCK>(laneB, lamB, j, scrB, load_rec(drawbuf + (groupBase + kk) * 4));
                        wave_sync_lds();
                    }
                }
            }
            wave_sync_lds();
            if (p.numDirs) {
                // this path's dE/d(initial segments) (scaled by w f) along the directions
                float x[kMaxDirs];
#pragma unroll
                for (uint32_t d = 0; d < kMaxDirs; ++d) x[d] = 0.0f;
                for (uint32_t s = 0; s < S; ++s) {
                    const uint32_t i = s * G + j;
                    if (i >= N) break;
                    const float l0 = smem[lamOff + s * SLOT + lane];
                    const float l1 = smem[lamOff + s * SLOT + BLOCK + lane];
                    const float l2 = smem[lamOff + s * SLOT + 2 * BLOCK + lane];
#pragma unroll
                    for (uint32_t d = 0; d < kMaxDirs; ++d) {
                        if (d < p.numDirs) {
                            const float *v = p.dirs + (uint64_t)d * 3u * N + 3u * i;
                            x[d] = fmaf(l0, v[0], fmaf(l1, v[1], fmaf(l2, v[2], x[d])));
                        }
                    }
                }
#pragma unroll
                for (uint32_t d = 0; d < kMaxDirs; ++d) {
                    if constexpr (G > 1) x[d] = group_sum<G>(x[d]);
                }
                if (j == 0) {
                    const float q = 1.0f / factor;
#pragma unroll
                    for (uint32_t d = 0; d < kMaxDirs; ++d) {
                        const float x2 = x[d] * x[d];
                        st[3 * d] = fmaf(x2, 1.0f - q, st[3 * d]);
                        st[3 * d + 1] = fmaf(x2, q, st[3 * d + 1]);
                        st[3 * d + 2] = fmaf(x[d], w0, st[3 * d + 2]);
                    }
                    st[3 * kMaxDirs] = fmaf(w0 * w0, factor, st[3 * kMaxDirs]);
                }
            }
            for (uint32_t s = 0; s < S; ++s) {
                const uint32_t i = s * G + j;
                if (i < N) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) atomicAdd(&gradS[3 * i + c], smem[lamOff + s * SLOT + c * BLOCK + lane]);
                }
            }
        }
        wave_sync_lds();  // neighbours' columns are read before the next path is loaded
    }
    __syncthreads();
    if (p.tableGrad) {
        float *gt = p.gradTablePartial + (uint64_t)blockIdx.x * p.tableSize;
        for (int i = threadIdx.x; i < p.tableSize; i += BLOCK) gt[i] = gradT[i];
    }
    if constexpr (SWEEP) {
        float *gs = p.gradSegPartial + (uint64_t)blockIdx.x * (3 * N);
        for (uint32_t i = threadIdx.x; i < 3 * N; i += BLOCK) gs[i] = gradS[i];
        if (p.numDirs) {
            // block sum of the leaders' statistics: wave sums, then one row per
            // wave in LDS (the state rows are free after the barrier above)
#pragma unroll
            for (uint32_t k = 0; k < kDirStatsMax; ++k) {
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) st[k] += __shfl_xor(st[k], o, 64);
            }
            const uint32_t wave = threadIdx.x >> 6;
            if ((threadIdx.x & 63u) == 0) {
#pragma unroll
                for (uint32_t k = 0; k < kDirStatsMax; ++k) smem[wave * kDirStatsMax + k] = st[k];
            }
            __syncthreads();
            if (threadIdx.x < kDirStatsMax) {
                float v = 0.0f;
                for (uint32_t w = 0; w < BLOCK / 64; ++w) v += smem[w * kDirStatsMax + threadIdx.x];
                p.dirStats[(uint64_t)blockIdx.x * kDirStatsMax + threadIdx.x] = v;
            }
        }
    }
}

// Each block takes 4096 consecutive paths (16 per thread, coalesced), keeps
// its paths' decisions in registers, reserves its slice of the list with one
// atomic and writes the kept paths there.
constexpr uint32_t kPickPerThread = 16;

__global__ void __launch_bounds__(256) adjoint_pick_kernel(const float2 *pathLog2, const FinalOut *fin,
      uint64_t numPaths, uint64_t pathOffset, uint64_t seed, float kappa, AdjPick *pick, uint32_t *count)
{
    __shared__ uint32_t waveTotal[4];
    __shared__ uint32_t blockBase;
    const float M = (float)fin->maxLog2;
    const bool finiteM = M > -INFINITY && M < INFINITY;
    // q_p = w_p * scale, scale = numPaths / (kappa S)
    const double scale = (double)numPaths / ((double)kappa * fin->sumScaled);
    const uint64_t first = (uint64_t)blockIdx.x * (256u * kPickPerThread);
    uint32_t keepBits = 0;
    float factor[kPickPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kPickPerThread; ++k) {
        factor[k] = 0.0f;
        const uint64_t local = first + k * 256u + threadIdx.x;
        if (local >= numPaths) continue;
        const float2 lw = pathLog2[local];
        const float wp = (finiteM && lw.x > -INFINITY && lw.x < INFINITY) ? exp2f((lw.x - M) + lw.y) : 0.0f;
        if (!(wp > 0.0f)) continue;
        const double q = (double)wp * scale;
        if (q >= 1.0) {
            keepBits |= 1u << k;
            factor[k] = 1.0f;
            continue;
        }
        const uint64_t path = pathOffset + local;
        const U4 r = philox4x32_10(0xA0000000u, 0u, (uint32_t)path, (uint32_t)(path >> 32), (uint32_t)seed,
              (uint32_t)(seed >> 32));
        if ((double)(r.x >> 8) * 0x1p-24 < q) {
            keepBits |= 1u << k;
            factor[k] = (float)(1.0 / q);
        }
    }
    // exclusive prefix of the kept counts over the block
    const uint32_t mine = (uint32_t)__popc(keepBits);
    const uint32_t wl = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (wl >= (uint32_t)o) incl += v;
    }
    if (wl == 63) waveTotal[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = waveTotal[0] + waveTotal[1] + waveTotal[2] + waveTotal[3];
        blockBase = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    uint32_t slot = blockBase + incl - mine;
    for (uint32_t w = 0; w < wave; ++w) slot += waveTotal[w];
#pragma unroll
    for (uint32_t k = 0; k < kPickPerThread; ++k) {
        if (keepBits & (1u << k)) {
            pick[slot].local = first + k * 256u + threadIdx.x;
            pick[slot].factor = factor[k];
            pick[slot].pad = 0u;
            ++slot;
        }
    }
}

__global__ void __launch_bounds__(256) grad_reduce_kernel(const float *partial, uint32_t numBlocks,
      uint32_t numEntries, const FinalOut *fin, double *out)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= numEntries) return;
    FF acc { 0.0f, 0.0f };
    for (uint32_t b = 0; b < numBlocks; ++b) acc = ff_add_f(acc, partial[(uint64_t)b * numEntries + e]);
    const double S = fin->sumScaled;
    const double ln10 = 2.30258509299404568402;
    out[e] = (S > 0.0) ? ((double)acc.hi + (double)acc.lo) / (S * ln10) : 0.0;
}

#define TWISTY_ADJ_INSTANCES(X)                                                                                   \
    X(1, 128) X(2, 256) X(4, 256) X(4, 512) X(8, 256) X(8, 512) X(16, 512) X(16, 1024) X(32, 512) X(32, 1024)    \
    X(64, 256) X(64, 512) X(64, 1024)

template<int G, int B, bool SW>
static hipError_t launch_adj_gb(const AdjParams &p, uint32_t grid, uint32_t lds, hipStream_t s)
{
    auto *fn = adjoint_kernel<G, B, SW>;
    hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, s, p);
    return hipGetLastError();
}

int adjoint_block_sizes(uint32_t G, bool, uint32_t *out, int cap)
{
    int n = 0;
#define X(g, b)                                                                                                   \
    if (G == g && n < cap) out[n++] = b;
    TWISTY_ADJ_INSTANCES(X)
#undef X
    return n;
}

hipError_t launch_adjoint(const AdjParams &p, uint32_t G, bool sweep, uint32_t grid, uint32_t block, uint32_t lds,
      hipStream_t s)
{
#define X(g, b)                                                                                                   \
    if (G == g && block == b) return sweep ? launch_adj_gb<g, b, true>(p, grid, lds, s)                         \
                                           : launch_adj_gb<g, b, false>(p, grid, lds, s);
    TWISTY_ADJ_INSTANCES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_adjoint_pick(const float2 *pathLog2, const FinalOut *fin, uint64_t numPaths, uint64_t pathOffset,
      uint64_t seed, float kappa, AdjPick *pick, uint32_t *count, hipStream_t s)
{
    const uint64_t per = 256u * kPickPerThread;
    const uint64_t blocks = (numPaths + per - 1) / per;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adjoint_pick_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, pathLog2, fin, numPaths, pathOffset,
          seed, kappa, pick, count);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const float *partial, uint32_t numBlocks, uint32_t numEntries, const FinalOut *fin,
      double *out, hipStream_t s)
{
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((numEntries + 255) / 256), dim3(256), 0, s, partial, numBlocks,
          numEntries, fin, out);
    return hipGetLastError();
}

}  // namespace twisty_hip
