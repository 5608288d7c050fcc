// Fused D3Q19 pull-stream + collide for gfx950: the kernel templates shared
// by bf_fluid_fused.hip, bf_fluid_strict.hip and bf_fluid.hip.
//
// Replaces the fluid part of stepFluidTRT (Sources/BirdFlowMetal/Metal/
// BirdFlow.metal:2507-3236).  The lattice is split into 64x4 tiles, one
// 256-lane workgroup each:
//   * fluid_fast    — tiles with no solid cell (current or previous step) in
//                     their 1-cell halo: pure pull stream with per-face
//                     boundary sources, collision, sponge blend, store.
//   * fluid_general — the listed tiles: solid nodes, refill of uncovered
//                     nodes, halfway or interpolated moving bounce-back,
//                     momentum-exchange loads with cover/uncover impulses,
//                     sponge.  Per-cell FP32 loads are summed in FP64 with
//                     wavefront shuffles and an LDS tree into one partial per
//                     tile; reduce_loads_* finishes the fixed-order sum.
// Two arithmetic forms of the collision, chosen per context:
//   * ARITH_STRICT — the reference's expressions in the reference's order,
//                    no contraction: bit-identical to the FP32 oracle.
//   * ARITH_FUSED  — the same operators regrouped over opposite-direction
//                    pairs with fused multiply-adds (about a third of the
//                    instructions); differs from STRICT by rounding only.
// Both kernels share the collision code; a fluid cell gets bit-identical
// results from either.
#pragma once

#include "bf_launch.h"

// Operand list of an empty asm that pins 19 values in registers: every load
// feeding them is issued before any is consumed.
#define BF_ASM19(a)                                                                                           \
    "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]),          \
        "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15]), \
        "+v"(a[16]), "+v"(a[17]), "+v"(a[18])

namespace bf {

// ---------------------------------------------------------------------------
// Collision (BirdFlow.metal:2888-3174).  f: post-stream populations; writes
// post[] (before sponge).  act/restr: positivity-limiter activation and
// maximum restriction.
// ---------------------------------------------------------------------------
template <int q>
__device__ __forceinline__ float dir_dot(float ux, float uy, float uz) {
    // c_q . u with the zero terms dropped; identical in value to the full
    // three-term dot product because c has entries 0/+-1.
    float r = 0.0f;
    bool first = true;
    if constexpr (cx(q) != 0) { r = cx(q) > 0 ? ux : -ux; first = false; }
    if constexpr (cy(q) != 0) { r = first ? (cy(q) > 0 ? uy : -uy) : (cy(q) > 0 ? r + uy : r - uy); first = false; }
    if constexpr (cz(q) != 0) { r = first ? (cz(q) > 0 ? uz : -uz) : (cz(q) > 0 ? r + uz : r - uz); first = false; }
    return r;
}

template <int q>
__device__ __forceinline__ float feq_q(float wrho, float u2_15, float ux, float uy, float uz) {
    float cu = dir_dot<q>(ux, uy, uz);
    return wrho * (1.0f + 3.0f * cu + 4.5f * cu * cu - u2_15);
}

// Wavefront issue priority (s_setprio takes an immediate).
__device__ __forceinline__ void set_priority(int level) {
    switch (level) {
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        case 3: __builtin_amdgcn_s_setprio(3); break;
        default: break;
    }
}

template <int... Q>
struct QSeq {};

template <typename F, int... Q>
__device__ __forceinline__ void for_q(QSeq<Q...>, F&& fn) {
    (fn(std::integral_constant<int, Q>{}), ...);
}
using AllQ = QSeq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18>;

__device__ __forceinline__ void moments(const float (&f)[19], float& rho, float& mx, float& my, float& mz) {
    // sum in ascending q, the order of the reference loop
    rho = 0.0f;
    mx = 0.0f;
    my = 0.0f;
    mz = 0.0f;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        rho = rho + f[q];
        if constexpr (cx(q) > 0) mx = mx + f[q];
        if constexpr (cx(q) < 0) mx = mx - f[q];
        if constexpr (cy(q) > 0) my = my + f[q];
        if constexpr (cy(q) < 0) my = my - f[q];
        if constexpr (cz(q) > 0) mz = mz + f[q];
        if constexpr (cz(q) < 0) mz = mz - f[q];
    });
}

template <int COLL>
__device__ __forceinline__ void collide(const float (&f)[19], float rho, float ux, float uy, float uz,
                                        float op, float om, float (&post)[19], float& act, float& restr) {
    float feq[19];
    float u2_15 = 1.5f * (ux * ux + uy * uy + uz * uz);
    float w0 = weight<float>(0) * rho, w1 = weight<float>(1) * rho, w2 = weight<float>(7) * rho;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        feq[q] = feq_q<q>(q == 0 ? w0 : (q <= 6 ? w1 : w2), u2_15, ux, uy, uz);
    });
    act = 0.0f;
    restr = 0.0f;
    if constexpr (COLL == COLL_TRT) {
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int o = opp(q);
            float fp = 0.5f * (f[q] + f[o]);
            float fm = 0.5f * (f[q] - f[o]);
            float ep = 0.5f * (feq[q] + feq[o]);
            float em = 0.5f * (feq[q] - feq[o]);
            post[q] = f[q] - op * (fp - ep) - om * (fm - em);
        });
    } else if constexpr (COLL == COLL_TRT_SYMMETRIC_LIMITER) {
        float sinc[19], ainc[19];
        float sscale = 1.0f;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int o = opp(q);
            float fp = 0.5f * (f[q] + f[o]);
            float fm = 0.5f * (f[q] - f[o]);
            float ep = 0.5f * (feq[q] + feq[o]);
            float em = 0.5f * (feq[q] - feq[o]);
            float si = -op * (fp - ep);
            float ai = -om * (fm - em);
            sinc[q] = si;
            ainc[q] = ai;
            if (si < 0.0f) {
                float base = f[q] + ai;
                float floorv = fmaxf(1.0e-12f, 1.0e-6f * fmaxf(feq[q], 0.0f));
                if (base + si < floorv) {
                    float cand = (base - floorv) / fmaxf(-si, 1.0e-30f);
                    cand = fminf(fmaxf(cand, 0.0f), 1.0f);
                    sscale = fminf(sscale, cand);
                }
            }
        });
        if (sscale < 1.0f) {
            act = 1.0f;
            restr = 1.0f - sscale;
        }
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int o = opp(q);
            if (sscale < 1.0f) {
                post[q] = f[q] + sscale * sinc[q] + ainc[q];
            } else {
                float fp = 0.5f * (f[q] + f[o]);
                float fm = 0.5f * (f[q] - f[o]);
                float ep = 0.5f * (feq[q] + feq[o]);
                float em = 0.5f * (feq[q] - feq[o]);
                post[q] = f[q] - op * (fp - ep) - om * (fm - em);
            }
        });
    } else {
        // Regularized / RR3 (BirdFlow.metal:2917-3062).  Terms whose factor is
        // a zero component of c_q are dropped at compile time: each is an
        // exact +-0 that leaves the running sum unchanged, so the result is
        // bit-identical to the full expression while the instruction count
        // falls by about half.
        constexpr bool RR3 = COLL == COLL_RR3;
        const float c2 = cs2<float>();
        const float h_one = 1.0f - c2;   // c_a c_a - cs2 for c_a = +-1
        const float h_zero = 0.0f - c2;  // and for c_a = 0
        float dg0 = 0.0f, dg1 = 0.0f, dg2 = 0.0f, od0 = 0.0f, od1 = 0.0f, od2 = 0.0f;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const float neq = f[q] - feq[q];
            dg0 = dg0 + neq * (cx(q) != 0 ? h_one : h_zero);
            dg1 = dg1 + neq * (cy(q) != 0 ? h_one : h_zero);
            dg2 = dg2 + neq * (cz(q) != 0 ? h_one : h_zero);
            if constexpr (cx(q) * cy(q) > 0) od0 = od0 + neq;
            if constexpr (cx(q) * cy(q) < 0) od0 = od0 - neq;
            if constexpr (cx(q) * cz(q) > 0) od1 = od1 + neq;
            if constexpr (cx(q) * cz(q) < 0) od1 = od1 - neq;
            if constexpr (cy(q) * cz(q) > 0) od2 = od2 + neq;
            if constexpr (cy(q) * cz(q) < 0) od2 = od2 - neq;
        });
        float aXxy = 2.0f * ux * od0 + uy * dg0;
        float aXxz = 2.0f * ux * od1 + uz * dg0;
        float aXyy = ux * dg1 + 2.0f * uy * od0;
        float aXzz = ux * dg2 + 2.0f * uz * od1;
        float aYyz = 2.0f * uy * od2 + uz * dg1;
        float aYzz = uy * dg2 + 2.0f * uz * od2;
        float unb[19];
        float scale = 1.0f;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int X = cx(q), Y = cy(q), Z = cz(q);
            const float hxx = X != 0 ? h_one : h_zero, hyy = Y != 0 ? h_one : h_zero, hzz = Z != 0 ? h_one : h_zero;
            float contraction = hxx * dg0 + hyy * dg1 + hzz * dg2;
            // at most one off-diagonal product c_a c_b is nonzero
            if constexpr (X * Y != 0) contraction = contraction + 2.0f * (X * Y > 0 ? od0 : -od0);
            if constexpr (X * Z != 0) contraction = contraction + 2.0f * (X * Z > 0 ? od1 : -od1);
            if constexpr (Y * Z != 0) contraction = contraction + 2.0f * (Y * Z > 0 ? od2 : -od2);
            const float w = weight<float>(q);
            float reg = w * 4.5f * contraction;
            if constexpr (RR3 && q != 0) {
                // sum in the reference's order of the terms whose leading
                // velocity component is nonzero
                float third = 0.0f;
                bool any = false;
                auto add = [&](float t) { third = any ? third + t : t; any = true; };
                if constexpr (Y != 0) add((Y > 0 ? hxx : -hxx) * aXxy);
                if constexpr (Z != 0) add((Z > 0 ? hxx : -hxx) * aXxz);
                if constexpr (X != 0) add((X > 0 ? hyy : -hyy) * aXyy);
                if constexpr (X != 0) add((X > 0 ? hzz : -hzz) * aXzz);
                if constexpr (Z != 0) add((Z > 0 ? hyy : -hyy) * aYyz);
                if constexpr (Y != 0) add((Y > 0 ? hzz : -hzz) * aYzz);
                reg = reg + w * 13.5f * third;
            }
            const float cand = feq[q] + (1.0f - op) * reg;
            unb[q] = cand;
            const float floorv = fmaxf(1.0e-12f, 1.0e-6f * fmaxf(feq[q], 0.0f));
            if (cand < floorv) {
                float adm = (feq[q] - floorv) / fmaxf(feq[q] - cand, 1.0e-30f);
                adm = fminf(fmaxf(adm, 0.0f), 1.0f);
                scale = fminf(scale, adm);
            }
        });
        if (scale < 1.0f) {
            act = 1.0f;
            restr = 1.0f - scale;
        }
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            post[q] = feq[q] + scale * (unb[q] - feq[q]);
        });
    }
}

// ---------------------------------------------------------------------------
// Fused arithmetic (ARITH_FUSED).  The D3Q19 directions pair up as (q, q+1)
// for odd q, with c_{q+1} = -c_q.  A pair enters every even moment (density,
// second moments, the even part of the equilibrium) through f_q + f_{q+1}
// and every odd one (momentum, the odd equilibrium part, the third-order
// Hermite terms) through f_q - f_{q+1}, so each operator of collide<COLL>
// yields an even part P and an odd part Q per pair with post_q = P + Q and
// post_{q+1} = P - Q.  Products are fused into FMAs.  The positivity limiters
// run their per-direction form only in cells where some post-collision
// population falls below 1e-12 + 1e-6 |feq|; elsewhere they cannot act.
// ---------------------------------------------------------------------------
using AllPairs = QSeq<0, 1, 2, 3, 4, 5, 6, 7, 8>;   // pair k: q = 2k + 1 and q + 1

template <int COLL>
__device__ __forceinline__ void collide_fused(const float (&f)[19], float op, float om, float (&post)[19],
                                              float& act, float& restr) {
    // -0.0f starts each sum so that the first addition is exact.
    float sp[9], dp[9];
    float rho = f[0], mx = -0.0f, my = -0.0f, mz = -0.0f;
    for_q(AllPairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        sp[k] = f[q] + f[q + 1];
        dp[k] = f[q] - f[q + 1];
        rho = rho + sp[k];
        if constexpr (cx(q) > 0) mx = mx + dp[k];
        if constexpr (cx(q) < 0) mx = mx - dp[k];
        if constexpr (cy(q) > 0) my = my + dp[k];
        if constexpr (cy(q) < 0) my = my - dp[k];
        if constexpr (cz(q) > 0) mz = mz + dp[k];
        if constexpr (cz(q) < 0) mz = mz - dp[k];
    });
    rho = fmaxf(rho, 1.0e-8f);
    const float inv = 1.0f / rho;
    const float ux = mx * inv, uy = my * inv, uz = mz * inv;
    const float base = fmaf(-1.5f, fmaf(ux, ux, fmaf(uy, uy, uz * uz)), 1.0f);
    // equilibrium of pair k: feq_q = A + B, feq_{q+1} = A - B
    float A[9], B[9];
    for_q(AllPairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        const float wr = rho * weight<float>(q);
        const float cu = dir_dot<q>(ux, uy, uz);
        A[k] = wr * fmaf(4.5f * cu, cu, base);
        B[k] = (3.0f * wr) * cu;
    });
    const float feq0 = (rho * weight<float>(0)) * base;
    act = 0.0f;
    restr = 0.0f;

    if constexpr (COLL == COLL_TRT || COLL == COLL_TRT_SYMMETRIC_LIMITER) {
        // post = f - omega+ (f^s - feq^s) - omega- (f^a - feq^a)
        float hs[9], ha[9];
        post[0] = fmaf(-op, f[0] - feq0, f[0]);
        for_q(AllPairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            hs[k] = fmaf(0.5f, sp[k], -A[k]);
            ha[k] = fmaf(0.5f, dp[k], -B[k]);
            post[q] = fmaf(-om, ha[k], fmaf(-op, hs[k], f[q]));
            post[q + 1] = fmaf(om, ha[k], fmaf(-op, hs[k], f[q + 1]));
        });
        if constexpr (COLL == COLL_TRT_SYMMETRIC_LIMITER) {
            float lo = fmaf(-1.0e-6f, fabsf(feq0), post[0]);
            for_q(AllPairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                lo = fminf(lo, fminf(fmaf(-1.0e-6f, fabsf(A[k] + B[k]), post[q]),
                                     fmaf(-1.0e-6f, fabsf(A[k] - B[k]), post[q + 1])));
            });
            if (lo < 2.0e-12f) {
                // scale the symmetric increment so that no population falls
                // below its floor (BirdFlow.metal:3100-3140)
                float sscale = 1.0f;
                auto limit = [&](float fq, float feq, float si, float ai) {
                    if (si < 0.0f) {
                        const float b = fq + ai;
                        const float floorv = fmaxf(1.0e-12f, 1.0e-6f * fmaxf(feq, 0.0f));
                        if (b + si < floorv) {
                            const float cand = (b - floorv) / fmaxf(-si, 1.0e-30f);
                            sscale = fminf(sscale, fminf(fmaxf(cand, 0.0f), 1.0f));
                        }
                    }
                };
                limit(f[0], feq0, -op * (f[0] - feq0), 0.0f);
                for_q(AllPairs{}, [&](auto kc) {
                    constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                    limit(f[q], A[k] + B[k], -op * hs[k], -om * ha[k]);
                    limit(f[q + 1], A[k] - B[k], -op * hs[k], om * ha[k]);
                });
                if (sscale < 1.0f) {
                    act = 1.0f;
                    restr = 1.0f - sscale;
                    post[0] = fmaf(sscale, -op * (f[0] - feq0), f[0]);
                    for_q(AllPairs{}, [&](auto kc) {
                        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                        post[q] = fmaf(sscale, -op * hs[k], f[q]) - om * ha[k];
                        post[q + 1] = fmaf(sscale, -op * hs[k], f[q + 1]) + om * ha[k];
                    });
                }
            }
        }
    } else {
        // Regularized / RR3 (BirdFlow.metal:2917-3062): non-equilibrium
        // second moments from the pair sums, Hermite reconstruction.
        constexpr bool RR3 = COLL == COLL_RR3;
        const float c2 = cs2<float>();
        const float h_one = 1.0f - c2;
        float nk[9];
        float S = f[0] - feq0, Px = -0.0f, Py = -0.0f, Pz = -0.0f;
        float od0 = -0.0f, od1 = -0.0f, od2 = -0.0f;
        for_q(AllPairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            nk[k] = fmaf(-2.0f, A[k], sp[k]);
            S = S + nk[k];
            if constexpr (cx(q) != 0) Px = Px + nk[k];
            if constexpr (cy(q) != 0) Py = Py + nk[k];
            if constexpr (cz(q) != 0) Pz = Pz + nk[k];
            if constexpr (cx(q) * cy(q) > 0) od0 = od0 + nk[k];
            if constexpr (cx(q) * cy(q) < 0) od0 = od0 - nk[k];
            if constexpr (cx(q) * cz(q) > 0) od1 = od1 + nk[k];
            if constexpr (cx(q) * cz(q) < 0) od1 = od1 - nk[k];
            if constexpr (cy(q) * cz(q) > 0) od2 = od2 + nk[k];
            if constexpr (cy(q) * cz(q) < 0) od2 = od2 - nk[k];
        });
        // diagonal moments of (c_a c_a - cs2) and their trace term
        const float dg0 = fmaf(-c2, S, Px), dg1 = fmaf(-c2, S, Py), dg2 = fmaf(-c2, S, Pz);
        const float c2T = c2 * ((dg0 + dg1) + dg2);
        const float om1 = 1.0f - op;
        float aXxy = 0.0f, aXxz = 0.0f, aXyy = 0.0f, aXzz = 0.0f, aYyz = 0.0f, aYzz = 0.0f;
        if constexpr (RR3) {
            const float ux2 = 2.0f * ux, uy2 = 2.0f * uy, uz2 = 2.0f * uz;
            aXxy = fmaf(ux2, od0, uy * dg0);
            aXxz = fmaf(ux2, od1, uz * dg0);
            aXyy = fmaf(uy2, od0, ux * dg1);
            aXzz = fmaf(uz2, od1, ux * dg2);
            aYyz = fmaf(uy2, od2, uz * dg1);
            aYzz = fmaf(uz2, od2, uy * dg2);
        }
        post[0] = fmaf(-4.5f * weight<float>(0) * om1, c2T, feq0);
        for_q(AllPairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            constexpr int X = cx(q), Y = cy(q), Z = cz(q);
            // H2 : Pi_neq = sum over the nonzero axes of dg_a - cs2 trace
            // + 2 c_a c_b od_ab
            float contr = -c2T;
            if constexpr (X != 0) contr = contr + dg0;
            if constexpr (Y != 0) contr = contr + dg1;
            if constexpr (Z != 0) contr = contr + dg2;
            if constexpr (X * Y != 0) contr = fmaf(X * Y > 0 ? 2.0f : -2.0f, od0, contr);
            if constexpr (X * Z != 0) contr = fmaf(X * Z > 0 ? 2.0f : -2.0f, od1, contr);
            if constexpr (Y * Z != 0) contr = fmaf(Y * Z > 0 ? 2.0f : -2.0f, od2, contr);
            const float w = weight<float>(q);
            const float P = fmaf((4.5f * w) * om1, contr, A[k]);
            float Q = B[k];
            if constexpr (RR3) {
                // H3 terms c_lead h_other a: h = 1 - cs2 on a nonzero axis,
                // -cs2 on a zero one; split into the two factors.
                float t_one = -0.0f, t_zero = -0.0f;
                auto term = [&](auto lead, auto other_nonzero, float a) {
                    constexpr int Ld = decltype(lead)::value;
                    if constexpr (Ld != 0) {
                        const float v = Ld > 0 ? a : -a;
                        if constexpr (decltype(other_nonzero)::value) t_one = t_one + v;
                        else t_zero = t_zero + v;
                    }
                };
                using IY = std::integral_constant<int, Y>;
                using IZ = std::integral_constant<int, Z>;
                using IX = std::integral_constant<int, X>;
                using NX = std::integral_constant<bool, X != 0>;
                using NY = std::integral_constant<bool, Y != 0>;
                using NZ = std::integral_constant<bool, Z != 0>;
                term(IY{}, NX{}, aXxy);
                term(IZ{}, NX{}, aXxz);
                term(IX{}, NY{}, aXyy);
                term(IX{}, NZ{}, aXzz);
                term(IZ{}, NY{}, aYyz);
                term(IY{}, NZ{}, aYzz);
                const float third = fmaf(h_one, t_one, -c2 * t_zero);
                Q = fmaf((13.5f * w) * om1, third, Q);
            }
            post[q] = P + Q;
            post[q + 1] = P - Q;
        });
        float lo = fmaf(-1.0e-6f, fabsf(feq0), post[0]);
        for_q(AllPairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            lo = fminf(lo, fminf(fmaf(-1.0e-6f, fabsf(A[k] + B[k]), post[q]),
                                 fmaf(-1.0e-6f, fabsf(A[k] - B[k]), post[q + 1])));
        });
        if (lo < 2.0e-12f) {
            // positivity limiter: blend toward equilibrium just enough
            float scale = 1.0f;
            auto limit = [&](float feq, float cand) {
                const float floorv = fmaxf(1.0e-12f, 1.0e-6f * fmaxf(feq, 0.0f));
                if (cand < floorv) {
                    const float adm = (feq - floorv) / fmaxf(feq - cand, 1.0e-30f);
                    scale = fminf(scale, fminf(fmaxf(adm, 0.0f), 1.0f));
                }
            };
            limit(feq0, post[0]);
            for_q(AllPairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                limit(A[k] + B[k], post[q]);
                limit(A[k] - B[k], post[q + 1]);
            });
            if (scale < 1.0f) {
                act = 1.0f;
                restr = 1.0f - scale;
                post[0] = fmaf(scale, post[0] - feq0, feq0);
                for_q(AllPairs{}, [&](auto kc) {
                    constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                    post[q] = fmaf(scale, post[q] - (A[k] + B[k]), A[k] + B[k]);
                    post[q + 1] = fmaf(scale, post[q + 1] - (A[k] - B[k]), A[k] - B[k]);
                });
            }
        }
    }
}

// Pull-source coordinate along one axis with the face treatment; sets far
// when the far-field equilibrium replaces the population.
__host__ __device__ __forceinline__ int shift_axis(int v, int d, int n, int bc_lo, int bc_hi, bool& far) {
    int s = v - d;
    if (s < 0) {
        if (bc_lo == BC_PERIODIC) s = n - 1;
        else if (bc_lo == BC_OUTLET) s = 0;
        else far = true;
    } else if (s >= n) {
        if (bc_hi == BC_PERIODIC) s = 0;
        else if (bc_hi == BC_OUTLET) s = n - 1;
        else far = true;
    }
    return s;
}

// Pull sources of the six axis directions with their far flags; the
// source of a direction combines them per axis.
struct Sources {
    int xs_p, xs_m, ys_p, ys_m, zs_p, zs_m;
    bool xm_far, xp_far, ym_far, yp_far, zm_far, zp_far;
};

__device__ __forceinline__ Sources sources(const StepParams& p, int x, int y, int z) {
    Sources s{};
    s.xm_far = s.xp_far = s.ym_far = s.yp_far = s.zm_far = s.zp_far = false;
    s.xs_p = shift_axis(x, 1, p.g.nx, p.bc[0][0], p.bc[0][1], s.xm_far);
    s.xs_m = shift_axis(x, -1, p.g.nx, p.bc[0][0], p.bc[0][1], s.xp_far);
    s.ys_p = shift_axis(y, 1, p.g.ny, p.bc[1][0], p.bc[1][1], s.ym_far);
    s.ys_m = shift_axis(y, -1, p.g.ny, p.bc[1][0], p.bc[1][1], s.yp_far);
    s.zs_p = shift_axis(z, 1, p.g.nz, p.bc[2][0], p.bc[2][1], s.zm_far);
    s.zs_m = shift_axis(z, -1, p.g.nz, p.bc[2][0], p.bc[2][1], s.zp_far);
    return s;
}

// Source cell of direction q: its x and row (y + ny z), and whether it lies
// beyond a far-field face.
template <int q>
__device__ __forceinline__ void source_xr(const Sources& s, const Grid& g, int x, int y, int z, int& xs,
                                          uint32_t& rows, bool& far) {
    xs = cx(q) > 0 ? s.xs_p : (cx(q) < 0 ? s.xs_m : x);
    const int ys = cy(q) > 0 ? s.ys_p : (cy(q) < 0 ? s.ys_m : y);
    const int zs = cz(q) > 0 ? s.zs_p : (cz(q) < 0 ? s.zs_m : z);
    far = (cx(q) > 0 && s.xm_far) || (cx(q) < 0 && s.xp_far) || (cy(q) > 0 && s.ym_far) || (cy(q) < 0 && s.yp_far) ||
          (cz(q) > 0 && s.zm_far) || (cz(q) < 0 && s.zp_far);
    rows = (uint32_t)ys + (uint32_t)g.ny * (uint32_t)zs;
}

template <int q>
__device__ __forceinline__ uint32_t source_of(const Sources& s, const Grid& g, int x, int y, int z, bool& far) {
    int xs;
    uint32_t rows;
    source_xr<q>(s, g, x, y, z, xs, rows, far);
    return (uint32_t)xs + (uint32_t)g.nx * rows;
}

// (x, q) is pulled by nobody when x + c_q leaves the domain through a
// far-field face
template <int q>
__device__ __forceinline__ bool orphan(const Sources& s) {
    return (cx(q) > 0 && s.xp_far) || (cx(q) < 0 && s.xm_far) || (cy(q) > 0 && s.yp_far) || (cy(q) < 0 && s.ym_far) ||
           (cz(q) > 0 && s.zp_far) || (cz(q) < 0 && s.zm_far);
}

template <int S>
__device__ __forceinline__ void write_post(typename Store<S>::T* __restrict__ fout, int x, uint32_t row,
                                           const float (&post)[19], float sponge, const StepParams& p,
                                           float& crho, float& cmx, float& cmy, float& cmz) {
    crho = 0.0f;
    cmx = 0.0f;
    cmy = 0.0f;
    cmz = 0.0f;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        float v = post[q];
        if (sponge > 0.0f) v = v + (p.far_eq[q] - v) * sponge;
        Store<S>::store(fout, pop_index<S>(p.g, q, x, row), q, v);
        crho = crho + v;
        if constexpr (cx(q) > 0) cmx = cmx + v;
        if constexpr (cx(q) < 0) cmx = cmx - v;
        if constexpr (cy(q) > 0) cmy = cmy + v;
        if constexpr (cy(q) < 0) cmy = cmy - v;
        if constexpr (cz(q) > 0) cmz = cmz + v;
        if constexpr (cz(q) < 0) cmz = cmz - v;
    });
}

__device__ __forceinline__ void capture(float* __restrict__ density, float4* __restrict__ velocity, uint32_t gid,
                                        float crho, float cmx, float cmy, float cmz) {
    crho = fmaxf(crho, 1.0e-8f);
    density[gid] = crho;
    velocity[gid] = make_float4(cmx / crho, cmy / crho, cmz / crho, 0.0f);
}

// ---------------------------------------------------------------------------
// Fast tiles.
// ---------------------------------------------------------------------------
template <int ARITH, int COLL, int S>
__global__ __launch_bounds__(256) void fluid_fast(const typename Store<S>::T* __restrict__ fin,
                                                  typename Store<S>::T* __restrict__ fout,
                                                  const uint8_t* __restrict__ static_flags,
                                                  const uint8_t* __restrict__ dyn_prev,
                                                  const uint8_t* __restrict__ dyn_cur,
                                                  float* __restrict__ density, float4* __restrict__ velocity,
                                                  uint32_t* __restrict__ activation, StepParams p) {
    set_priority(p.fast_priority);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    // Tiles with walls belong to fluid_general.  The flags are
    // loaded together with the populations and tested only before the
    // stores: a test up front would put a dependent memory round trip ahead
    // of every workgroup's population loads.
    const bool general_tile = (static_flags[tile] | dyn_prev[tile] | dyn_cur[tile]) != 0;
    int x0, y0, z;
    tile_origin(p.g, tile, x0, y0, z);
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    // Lanes of a partial tile beyond the grid load from the last column / row
    // (their tile is a general tile: nothing they load is used), so no load
    // leaves the lattice; in whole tiles the clamp changes nothing.
    const int x = min(x0 + (int)(threadIdx.x & (TILE_X - 1)), nx - 1);
    const int y = min(y0 + (int)(threadIdx.x / TILE_X), ny - 1);

    bool xm_far = false, xp_far = false, ym_far = false, yp_far = false, zm_far = false, zp_far = false;
    // source coordinates for c = +1 (s = v - 1) and c = -1 (s = v + 1)
    const int xs_p = shift_axis(x, 1, nx, p.bc[0][0], p.bc[0][1], xm_far);
    const int xs_m = shift_axis(x, -1, nx, p.bc[0][0], p.bc[0][1], xp_far);
    const int ys_p = shift_axis(y, 1, ny, p.bc[1][0], p.bc[1][1], ym_far);
    const int ys_m = shift_axis(y, -1, ny, p.bc[1][0], p.bc[1][1], yp_far);
    const int zs_p = shift_axis(z, 1, nz, p.bc[2][0], p.bc[2][1], zm_far);
    const int zs_m = shift_axis(z, -1, nz, p.bc[2][0], p.bc[2][1], zp_far);

    const uint32_t row_self = (uint32_t)y + (uint32_t)ny * (uint32_t)z;
    float f[19];
    bool far_q[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int xs = cx(q) > 0 ? xs_p : (cx(q) < 0 ? xs_m : x);
        const int ys = cy(q) > 0 ? ys_p : (cy(q) < 0 ? ys_m : y);
        const int zs = cz(q) > 0 ? zs_p : (cz(q) < 0 ? zs_m : z);
        const bool far = (cx(q) > 0 && xm_far) || (cx(q) < 0 && xp_far) || (cy(q) > 0 && ym_far)
            || (cy(q) < 0 && yp_far) || (cz(q) > 0 && zm_far) || (cz(q) < 0 && zp_far);
        const uint32_t row_src = (uint32_t)ys + (uint32_t)ny * (uint32_t)zs;
        // unconditional load: a far-field source reads an in-domain cell and
        // discards it
        f[q] = Store<S>::load(fin, far ? pop_index<S>(p.g, q, x, row_self) : pop_index<S>(p.g, q, xs, row_src), q);
        far_q[q] = far;
    });
    // One empty asm over all 19 loaded values: every load is issued, without
    // a branch, before any of them is consumed (otherwise the far-field
    // select turns each load into a conditional block, and the scheduler
    // starts the moment sums after a few loads and drains the queue), and
    // none sinks below the tile test.
    asm volatile("" : BF_ASM19(f));
    if (general_tile) return;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (far_q[q]) f[q] = p.far_eq[q];
    });

    float post[19];
    float act, restr;
    if constexpr (ARITH == ARITH_FUSED) {
        collide_fused<COLL>(f, p.omega_plus, p.omega_minus, post, act, restr);
    } else {
        float rho, mx, my, mz;
        moments(f, rho, mx, my, mz);
        rho = fmaxf(rho, 1.0e-8f);
        const float ux = mx / rho, uy = my / rho, uz = mz / rho;
        collide<COLL>(f, rho, ux, uy, uz, p.omega_plus, p.omega_minus, post, act, restr);
    }
    const uint32_t gid = (uint32_t)x + (uint32_t)nx * row_self;
    // wall-free sponge tiles stream here too (the sponge is a per-cell blend
    // toward the far-field equilibrium after the collision); a uniform branch,
    // so that without a sponge the store path carries no blend
    float crho, cmx, cmy, cmz;
    if (p.sponge_strength > 0.0f)
        write_post<S>(fout, x, row_self, post, sponge_factor(p, x, y, z), p, crho, cmx, cmy, cmz);
    else
        write_post<S>(fout, x, row_self, post, 0.0f, p, crho, cmx, cmy, cmz);
    guard_finite(p, crho);
    if (p.capture_fields) capture(density, velocity, gid, crho, cmx, cmy, cmz);
    if constexpr (COLL != COLL_TRT) {
        // Limiter activations of fast cells: an exact count and the maximum
        // restriction (order-independent integer and max atomics; restr >= 0
        // orders like its bit pattern).  Fast tiles are whole, so every lane
        // of the wavefront is active here.
        if (p.accumulate_loads) {
            const unsigned long long hit = __ballot(act != 0.0f);
            if (hit != 0) {
                float r = restr;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) r = fmaxf(r, __shfl_xor(r, off, 64));
                if ((threadIdx.x & 63) == 0) {
                    atomicAdd(&activation[0], (uint32_t)__popcll(hit));
                    atomicMax(&activation[1], __float_as_uint(r));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fast tiles in 2-byte storage (FP16 / BF16 offsets): two cells per lane.
//
// One 256-lane workgroup per pair of x-adjacent tiles (128 x 4 cells); each
// wavefront streams one 128-cell row, lane l the cells X0 + 2l and
// X0 + 2l + 1, so every population load and store moves a 4-byte pair (256
// bytes per wavefront instruction instead of 128) and a wavefront has twice
// the bytes in flight per load.  Loads stay aligned: a direction with
// c_x = +1 needs the pair (X0 + 2l - 1, X0 + 2l), whose first half is the
// second half of the lane below's aligned pair (a DPP wavefront shift) and,
// for lane 0, the element at X0 - 1, a wave-uniform address (one scalar
// load); c_x = -1 mirrors this with the lane above and X0 + 128.  y / z
// sources and their face treatment are wave-uniform.  Each cell's arithmetic
// is exactly fluid_fast's, so a cell gets bit-identical results from either
// kernel.  Needs nx a multiple of 128 (whole tile pairs, rows starting at
// even elements; the launcher keeps fluid_fast otherwise); lanes of a tile
// on the general path load but do not store, and the rows of a partial tile
// beyond ny load from the last row.
// ---------------------------------------------------------------------------
template <int S>
struct PairStore;

template <>
struct PairStore<STORE_FP16_OFFSET> {
    __device__ static inline float lo(uint32_t u, int q) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu)) + weight<float>(q);
    }
    __device__ static inline float hi(uint32_t u, int q) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16)) + weight<float>(q);
    }
    __device__ static inline uint32_t pack(float a, float b, int q) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 h = {(_Float16)(a - weight<float>(q)), (_Float16)(b - weight<float>(q))};
        return __builtin_bit_cast(uint32_t, h);
    }
};

template <>
struct PairStore<STORE_BF16_OFFSET> {
    __device__ static inline float lo(uint32_t u, int q) { return __uint_as_float(u << 16) + weight<float>(q); }
    __device__ static inline float hi(uint32_t u, int q) {
        return __uint_as_float(u & 0xffff0000u) + weight<float>(q);
    }
    __device__ static inline uint32_t pack(float a, float b, int q) {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        const b2 h = {(__bf16)(a - weight<float>(q)), (__bf16)(b - weight<float>(q))};
        return __builtin_bit_cast(uint32_t, h);
    }
};

template <int ARITH, int COLL, int S>
__global__ __launch_bounds__(256) void fluid_fast_pair(const typename Store<S>::T* __restrict__ fin_t,
                                                       typename Store<S>::T* __restrict__ fout_t,
                                                       const uint8_t* __restrict__ static_flags,
                                                       const uint8_t* __restrict__ dyn_prev,
                                                       const uint8_t* __restrict__ dyn_cur,
                                                       float* __restrict__ density, float4* __restrict__ velocity,
                                                       uint32_t* __restrict__ activation, StepParams p) {
    static_assert(S == STORE_FP16_OFFSET || S == STORE_BF16_OFFSET, "two-byte storage only");
    set_priority(p.fast_priority);
    const uint32_t* __restrict__ fin = reinterpret_cast<const uint32_t*>(fin_t);
    uint32_t* __restrict__ fout = reinterpret_cast<uint32_t*>(fout_t);
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    const uint32_t pairs_x = (uint32_t)p.g.tiles_x / 2u;
    const uint32_t pair = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t per_plane = pairs_x * (uint32_t)p.g.tiles_y;
    const int z = (int)(pair / per_plane);
    const uint32_t r = pair - (uint32_t)z * per_plane;
    const uint32_t ty = r / pairs_x, px = r - ty * pairs_x;
    const uint32_t tile_a = ((uint32_t)z * (uint32_t)p.g.tiles_y + ty) * (uint32_t)p.g.tiles_x + 2u * px;
    // the six flag bytes load with the populations and are tested after them
    uint32_t flags_a = (uint32_t)static_flags[tile_a], flags_a1 = (uint32_t)dyn_prev[tile_a],
             flags_a2 = (uint32_t)dyn_cur[tile_a];
    uint32_t flags_b = (uint32_t)static_flags[tile_a + 1], flags_b1 = (uint32_t)dyn_prev[tile_a + 1],
             flags_b2 = (uint32_t)dyn_cur[tile_a + 1];
    const int lane = (int)(threadIdx.x & 63u);
    const int wave = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int X0 = (int)px * 2 * TILE_X;
    const int xa = X0 + 2 * lane;
    const int y = min((int)ty * TILE_Y + wave, ny - 1);

    // wave-uniform sources along y and z; the halo elements along x
    bool ym_far = false, yp_far = false, zm_far = false, zp_far = false, xlo_far = false, xhi_far = false;
    const int ys_p = shift_axis(y, 1, ny, p.bc[1][0], p.bc[1][1], ym_far);
    const int ys_m = shift_axis(y, -1, ny, p.bc[1][0], p.bc[1][1], yp_far);
    const int zs_p = shift_axis(z, 1, nz, p.bc[2][0], p.bc[2][1], zm_far);
    const int zs_m = shift_axis(z, -1, nz, p.bc[2][0], p.bc[2][1], zp_far);
    int xh_lo = shift_axis(X0, 1, nx, p.bc[0][0], p.bc[0][1], xlo_far);           // source of X0 along +x
    int xh_hi = shift_axis(X0 + 2 * TILE_X - 1, -1, nx, p.bc[0][0], p.bc[0][1], xhi_far);   // of X0 + 127 along -x
    if (xlo_far) xh_lo = X0;   // in-domain stand-in; the value is replaced
    if (xhi_far) xh_hi = X0;
    const uint32_t row_self = (uint32_t)y + (uint32_t)ny * (uint32_t)z;

    uint32_t v[19];
    uint32_t halo[19];         // the 4-byte word holding each x-direction's halo element
    bool far_yz[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int ys = cy(q) > 0 ? ys_p : (cy(q) < 0 ? ys_m : y);
        const int zs = cz(q) > 0 ? zs_p : (cz(q) < 0 ? zs_m : z);
        const bool far = (cy(q) > 0 && ym_far) || (cy(q) < 0 && yp_far) || (cz(q) > 0 && zm_far) ||
                         (cz(q) < 0 && zp_far);
        const uint32_t row_src = far ? row_self : (uint32_t)ys + (uint32_t)ny * (uint32_t)zs;
        const size_t base = pop_index<S>(p.g, q, 0, row_src);          // even: nx is even
        v[q] = fin[(base + (size_t)X0) / 2 + (size_t)lane];
        if constexpr (cx(q) != 0) halo[q] = fin[(base + (size_t)(cx(q) > 0 ? xh_lo : xh_hi)) / 2];
        far_yz[q] = far;
    });
    // Every population, halo and flag load issues before the first wait:
    // the empty asm statements consume all of them at once (otherwise the
    // compiler extracts each halo half right after its scalar load, waiting
    // on it, and tests the flags before the population loads).
    asm volatile("" : BF_ASM19(v), "+s"(halo[1]), "+s"(halo[2]), "+s"(halo[7]), "+s"(halo[8]), "+s"(halo[9]),
                 "+s"(halo[10]), "+s"(halo[11]), "+s"(halo[12]), "+s"(halo[13]), "+s"(halo[14]));
    asm volatile("" : "+v"(flags_a), "+v"(flags_a1), "+v"(flags_a2), "+v"(flags_b), "+v"(flags_b1), "+v"(flags_b2));
    const bool general_a = (flags_a | flags_a1 | flags_a2) != 0;
    const bool general_b = (flags_b | flags_b1 | flags_b2) != 0;
    if (general_a && general_b) return;
    // the halo half within its word (population rows start at even elements)
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (cx(q) != 0) {
            const int xh = cx(q) > 0 ? xh_lo : xh_hi;
            halo[q] = (xh & 1) ? (halo[q] >> 16) : (halo[q] & 0xffffu);
        }
    });

    // shift the x-neighbour halves into place
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (cx(q) > 0) {
            const uint32_t below = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[q], 0x138, 0xf, 0xf, false);
            const uint32_t first = lane == 0 ? halo[q] : (below >> 16);
            v[q] = first | (v[q] << 16);
        } else if constexpr (cx(q) < 0) {
            const uint32_t above = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[q], 0x130, 0xf, 0xf, false);
            const uint32_t second = lane == 63 ? halo[q] : (above & 0xffffu);
            v[q] = (v[q] >> 16) | (second << 16);
        }
    });
    const bool store = lane < 32 ? !general_a : !general_b;
    const bool sponge_on = p.sponge_strength > 0.0f;
    float act_n = 0.0f, restr_max = 0.0f;
    uint32_t out[19];
    float crho[2], cmx[2], cmy[2], cmz[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int x = xa + c;
        float f[19];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = c == 0 ? PairStore<S>::lo(v[q], q) : PairStore<S>::hi(v[q], q);
            bool far = far_yz[q];
            if constexpr (cx(q) > 0) far = far || (xlo_far && c == 0 && lane == 0);
            if constexpr (cx(q) < 0) far = far || (xhi_far && c == 1 && lane == 63);
            if (far) f[q] = p.far_eq[q];
        });
        float post[19];
        float act, restr;
        if constexpr (ARITH == ARITH_FUSED) {
            collide_fused<COLL>(f, p.omega_plus, p.omega_minus, post, act, restr);
        } else {
            float rho, mx, my, mz;
            moments(f, rho, mx, my, mz);
            rho = fmaxf(rho, 1.0e-8f);
            collide<COLL>(f, rho, mx / rho, my / rho, mz / rho, p.omega_plus, p.omega_minus, post, act, restr);
        }
        act_n += act;
        restr_max = fmaxf(restr_max, restr);
        const float sponge = sponge_on ? sponge_factor(p, x, y, z) : 0.0f;
        crho[c] = 0.0f;
        cmx[c] = 0.0f;
        cmy[c] = 0.0f;
        cmz[c] = 0.0f;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float val = post[q];
            if (sponge > 0.0f) val = val + (p.far_eq[q] - val) * sponge;
            if (c == 0) out[q] = __float_as_uint(val);
            else out[q] = PairStore<S>::pack(__uint_as_float(out[q]), val, q);
            crho[c] = crho[c] + val;
            if constexpr (cx(q) > 0) cmx[c] = cmx[c] + val;
            if constexpr (cx(q) < 0) cmx[c] = cmx[c] - val;
            if constexpr (cy(q) > 0) cmy[c] = cmy[c] + val;
            if constexpr (cy(q) < 0) cmy[c] = cmy[c] - val;
            if constexpr (cz(q) > 0) cmz[c] = cmz[c] + val;
            if constexpr (cz(q) < 0) cmz[c] = cmz[c] - val;
        });
    }
    if (store) {
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            __builtin_nontemporal_store(out[q], fout + (pop_index<S>(p.g, q, xa, row_self) >> 1));
        });
        guard_finite(p, crho[0] + crho[1]);
        if (p.capture_fields) {
            const uint32_t gid = (uint32_t)xa + (uint32_t)nx * row_self;
            capture(density, velocity, gid, crho[0], cmx[0], cmy[0], cmz[0]);
            capture(density, velocity, gid + 1, crho[1], cmx[1], cmy[1], cmz[1]);
        }
    }
    if constexpr (COLL != COLL_TRT) {
        // limiter activations of the stored cells: an exact count and the
        // maximum restriction (order-independent integer and max atomics)
        if (p.accumulate_loads) {
            const float hits = store ? act_n : 0.0f;
            const unsigned long long any = __ballot(hits != 0.0f);
            if (any != 0) {
                uint32_t count = (uint32_t)hits;
                float rm = store ? restr_max : 0.0f;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    count += (uint32_t)__shfl_xor((int)count, off, 64);
                    rm = fmaxf(rm, __shfl_xor(rm, off, 64));
                }
                if (lane == 0) {
                    atomicAdd(&activation[0], count);
                    atomicMax(&activation[1], __float_as_uint(rm));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// General tiles: full moving-boundary semantics of stepFluidTRT.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cross3(float ax, float ay, float az, float bx, float by, float bz,
                                       float& ox, float& oy, float& oz) {
    ox = ay * bz - az * by;
    oy = az * bx - ax * bz;
    oz = ax * by - ay * bx;
}

__device__ __forceinline__ bool inside(const Grid& g, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz;
}

__device__ __forceinline__ uint32_t flat(const Grid& g, int x, int y, int z) {
    return (uint32_t)x + (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
}

struct CellLoad {
    float fx, fy, fz, tx, ty, tz, act, restr;
};

template <int ARITH, int COLL, int S>
__device__ __forceinline__ void general_cell(const typename Store<S>::T* __restrict__ fin, typename Store<S>::T* __restrict__ fout,
                             const uint8_t* __restrict__ solid_prev, const uint8_t* __restrict__ solid_cur,
                             const float4* __restrict__ wall, float* __restrict__ density,
                             float4* __restrict__ velocity, const StepParams& p, int x, int y, int z,
                             float bx, float by, float bz, CellLoad& L) {
    const Grid& g = p.g;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t gid = (uint32_t)x + (uint32_t)g.nx * row;
    const float wx = p.origin[0] + ((float)x + 0.5f) * p.dx;
    const float wy = p.origin[1] + ((float)y + 0.5f) * p.dx;
    const float wz = p.origin[2] + ((float)z + 0.5f) * p.dx;
    const float ax = wx - bx, ay = wy - by, az = wz - bz;
    const bool was_solid = solid_prev[gid] != 0;
    const bool is_solid = solid_cur[gid] != 0;
    const int mode = p.force_mode;
    const bool include_link = p.load_component != 2;
    const bool include_topology = p.load_component != 1;
    const float fscale = p.force_scale;
    L = CellLoad{0, 0, 0, 0, 0, 0, 0, 0};

    if (is_solid) {
        const float4 wv = wall[gid];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            Store<S>::store(fout, pop_index<S>(g, q, x, row), q, equilibrium<float>(q, 1.0f, wv.x, wv.y, wv.z));
        });
        if (p.accumulate_loads && !was_solid && include_topology) {
            float prho = 0.0f, pmx = 0.0f, pmy = 0.0f, pmz = 0.0f;
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                float v = Store<S>::load(fin, pop_index<S>(g, q, x, row), q);
                prho = prho + v;
                pmx = pmx + v * (float)cx(q);
                pmy = pmy + v * (float)cy(q);
                pmz = pmz + v * (float)cz(q);
            });
            float cvx = mode == 6 ? pmx : pmx - prho * wv.x;
            float cvy = mode == 6 ? pmy : pmy - prho * wv.y;
            float cvz = mode == 6 ? pmz : pmz - prho * wv.z;
            L.fx = cvx * fscale;
            L.fy = cvy * fscale;
            L.fz = cvz * fscale;
            cross3(ax, ay, az, L.fx, L.fy, L.fz, L.tx, L.ty, L.tz);
        }
        if (p.capture_fields) {
            density[gid] = 1.0f;
            velocity[gid] = make_float4(wv.x, wv.y, wv.z, 0.0f);
        }
        return;
    }

    float f[19];
    float rho = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    float flx = 0.0f, fly = 0.0f, flz = 0.0f;
    float tqx = 0.0f, tqy = 0.0f, tqz = 0.0f;

    if (was_solid) {
        const float4 rv = wall[gid];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float v = equilibrium<float>(q, 1.0f, rv.x, rv.y, rv.z);
            f[q] = v;
            rho = rho + v;
            mx = mx + v * (float)cx(q);
            my = my + v * (float)cy(q);
            mz = mz + v * (float)cz(q);
        });
        if (p.accumulate_loads && include_topology && mode == 6) {
            float tx = -mx, ty = -my, tz = -mz;
            // two rounds of unconditional loads (a neighbour outside the
            // domain reads this cell and is ignored): the neighbours' part
            // ids, then the populations of the open directions
            auto nb_of = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int bx = x + cx(q), by = y + cy(q), bz = z + cz(q);
                return inside(g, bx, by, bz) ? flat(g, bx, by, bz) : gid;
            };
            // population opp(q) of that neighbour (of this cell outside the domain)
            auto nb_pop = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int bx = x + cx(q), by = y + cy(q), bz = z + cz(q);
                return inside(g, bx, by, bz) ? pop_index<S>(g, opp(q), bx, (uint32_t)by + (uint32_t)g.ny * (uint32_t)bz)
                                             : pop_index<S>(g, opp(q), x, row);
            };
            uint32_t mk[19];
            uint32_t inbits = 0;
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const uint32_t nb = nb_of(qc);
                mk[q] = (uint32_t)solid_prev[nb] | (uint32_t)solid_cur[nb];
                inbits |= (uint32_t)inside(g, x + cx(q), y + cy(q), z + cz(q)) << q;
            });
            asm volatile("" : BF_ASM19(mk));
            float own[19], far[19];
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                own[q] = Store<S>::load(fin, pop_index<S>(g, q, x, row), q);
                far[q] = Store<S>::load(fin, nb_pop(qc), opp(q));
            });
            asm volatile("" : BF_ASM19(own), BF_ASM19(far));
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (q != 0) {
                    if (((inbits >> q) & 1u) && mk[q] == 0) {
                        float s = own[q] + far[q];
                        tx = tx - s * (float)cx(q);
                        ty = ty - s * (float)cy(q);
                        tz = tz - s * (float)cz(q);
                    }
                }
            });
            flx = flx + tx;
            fly = fly + ty;
            flz = flz + tz;
            float ox, oy, oz;
            cross3(ax, ay, az, tx * fscale, ty * fscale, tz * fscale, ox, oy, oz);
            tqx = tqx + ox;
            tqy = tqy + oy;
            tqz = tqz + oz;
        }
    } else {
        const float mwd = p.far_rho;
        const bool interpolated_rule = p.wall_rule != WALL_HALFWAY;
        const float4 wg = wall[gid];
        // Loads in two passes.  First, unconditionally and without a branch
        // between them (a far-field source, or a far side outside the
        // domain, reads this cell and is ignored): every source's part id and
        // population, this cell's own populations (reflected and x-side
        // values of its links) and the part id of each link's far side.
        // Conditional loads here would cost one memory round trip per
        // direction.  Then, per link, the source's wall entry and the
        // far-side population.
        const Sources sq = sources(p, x, y, z);
        // source and far-side cell of direction q (far-field / outside: this
        // cell); recomputed where needed rather than kept in registers
        auto src_of = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            bool far;
            const uint32_t src = source_of<q>(sq, g, x, y, z, far);
            return far ? gid : src;
        };
        auto far_side_of = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
            return inside(g, fx, fy, fz) ? flat(g, fx, fy, fz) : gid;
        };
        // population opp(q) on the far side (of this cell outside the domain)
        auto far_side_pop = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
            return inside(g, fx, fy, fz) ? pop_index<S>(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz)
                                         : pop_index<S>(g, opp(q), x, row);
        };
        uint32_t partq[19], fpart[19];
        float fsrc[19], fown[19];
        uint32_t farbits = 0, okbits = 0;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            bool far;
            (void)source_of<q>(sq, g, x, y, z, far);
            farbits |= (uint32_t)far << q;
            okbits |= (uint32_t)inside(g, x + cx(q), y + cy(q), z + cz(q)) << q;
            const uint32_t src = src_of(qc), fid = far_side_of(qc);
            partq[q] = solid_cur[src];
            fpart[q] = solid_cur[fid];
            int xs;
            uint32_t rows;
            bool farq;
            source_xr<q>(sq, g, x, y, z, xs, rows, farq);
            fsrc[q] = Store<S>::load(fin, farq ? pop_index<S>(g, q, x, row) : pop_index<S>(g, q, xs, rows), q);
            fown[q] = Store<S>::load(fin, pop_index<S>(g, q, x, row), q);
        });
        asm volatile("" : BF_ASM19(fsrc), BF_ASM19(fown), BF_ASM19(partq), BF_ASM19(fpart));
        // links; far sides that force the halfway fallback; links whose part
        // counts toward the loads
        uint32_t linkbits = 0, fallbits = 0, selbits = 0;
        const int sel = p.selected_part;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            linkbits |= (uint32_t)(partq[q] != 0 && !((farbits >> q) & 1u)) << q;
            fallbits |= (uint32_t)(!((okbits >> q) & 1u) || fpart[q] != 0) << q;
            selbits |= (uint32_t)(sel == 0 || sel == (int)partq[q]) << q;
        });
        // Second load pass, six directions at a time: the wall entry of each link's
        // source and the population on its far side, all loaded before any is
        // used (a direction without a link reads this cell's entries and
        // ignores them); then the values of the group's directions in
        // ascending q.  A load per link inside the link branch would cost one
        // memory round trip per link direction of the wavefront.
        auto link_group = [&](auto seq, auto first_c, uint32_t gmask) {
            constexpr int q0 = decltype(first_c)::value;
            float4 wsg[7];
            float fog[7];
            for (int k = 0; k < 7; ++k) {
                wsg[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                fog[k] = 0.0f;
            }
            if (linkbits & gmask) {
                for_q(seq, [&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if constexpr (q != 0) {
                        const bool lk = (linkbits >> q) & 1u;
                        wsg[q - q0] = wall[lk ? src_of(qc) : gid];
                        fog[q - q0] = Store<S>::load(fin, lk ? far_side_pop(qc) : pop_index<S>(g, opp(q), x, row), opp(q));
                    }
                });
                asm volatile("" : "+v"(wsg[0].x), "+v"(wsg[0].y), "+v"(wsg[0].z), "+v"(wsg[0].w),
                             "+v"(wsg[1].x), "+v"(wsg[1].y), "+v"(wsg[1].z), "+v"(wsg[1].w),
                             "+v"(wsg[2].x), "+v"(wsg[2].y), "+v"(wsg[2].z), "+v"(wsg[2].w),
                             "+v"(wsg[3].x), "+v"(wsg[3].y), "+v"(wsg[3].z), "+v"(wsg[3].w),
                             "+v"(wsg[4].x), "+v"(wsg[4].y), "+v"(wsg[4].z), "+v"(wsg[4].w),
                             "+v"(wsg[5].x), "+v"(wsg[5].y), "+v"(wsg[5].z), "+v"(wsg[5].w),
                             "+v"(wsg[6].x), "+v"(wsg[6].y), "+v"(wsg[6].z), "+v"(wsg[6].w),
                             "+v"(fog[0]), "+v"(fog[1]), "+v"(fog[2]), "+v"(fog[3]), "+v"(fog[4]), "+v"(fog[5]),
                             "+v"(fog[6]));
            }
            for_q(seq, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr float dxq = (float)cx(q), dyq = (float)cy(q), dzq = (float)cz(q);
            float value;
            if ((farbits >> q) & 1u) {
                value = p.far_eq[q];
            } else {
                if ((linkbits >> q) & 1u) {
                    const float reflected = fown[opp(q)];
                    float lf = 0.5f;
                    const float4 ws = wsg[q - q0];
                    const float fo_far = fog[q - q0];
                    float wvx = ws.x, wvy = ws.y, wvz = ws.z;
                    bool interp = interpolated_rule;
                    if (interp) {
                        if (p.wall_rule == WALL_INTERPOLATED_SDF) {
                            // BirdFlow.metal:2062-2071: q from the fluid and
                            // solid signed distances of this link.
                            float df = fmaxf(wg.w, 0.0f);
                            float dsv = fminf(ws.w, 0.0f);
                            lf = fminf(fmaxf(df / fmaxf(df - dsv, 1.0e-6f), 1.0e-4f), 1.0f);
                        } else if (p.wall_rule == WALL_INTERPOLATED_LINKS) {
                            // BirdFlow.metal:2738-2745: the stored link fraction
                            const float v = p.links[((size_t)tile_of(g, x, y, z) * 19u + (size_t)q) * TILE_CELLS
                                                    + threadIdx.x];
                            lf = fminf(fmaxf(v, 1.0e-4f), 1.0f);
                        }
                        wvx = wg.x + (ws.x - wg.x) * lf;
                        wvy = wg.y + (ws.y - wg.y) * lf;
                        wvz = wg.z + (ws.z - wg.z) * lf;
                        if (lf <= 0.5f) {
                            if ((fallbits >> q) & 1u) {
                                interp = false;
                                lf = 0.5f;
                                wvx = ws.x;
                                wvy = ws.y;
                                wvz = ws.z;
                            }
                        }
                    }
                    const float dotdw = dxq * wvx + dyq * wvy + dzq * wvz;
                    const float corr = 2.0f * weight<float>(q) * mwd * dotdw / cs2<float>();
                    value = reflected + corr;
                    if (interp) {
                        if (lf <= 0.5f) {
                            const float fo = fo_far;
                            value = 2.0f * lf * reflected + (1.0f - 2.0f * lf) * fo + corr;
                        } else {
                            const float pin = fown[q];
                            value = (reflected + corr) / (2.0f * lf) + (2.0f * lf - 1.0f) * pin / (2.0f * lf);
                        }
                    }
                    if (p.accumulate_loads && include_link && ((selbits >> q) & 1u)) {
                        float lx = 0.0f, ly = 0.0f, lz = 0.0f;
                        if (mode == 0) {
                            lx = -(value * (dxq - wvx) - reflected * (-dxq - wvx));
                            ly = -(value * (dyq - wvy) - reflected * (-dyq - wvy));
                            lz = -(value * (dzq - wvz) - reflected * (-dzq - wvz));
                        } else if (mode == 1 || mode == 6) {
                            float s = -(value + reflected);
                            lx = s * dxq;
                            ly = s * dyq;
                            lz = s * dzq;
                        } else if (mode == 2) {
                            float s = -2.0f * reflected;
                            lx = s * dxq;
                            ly = s * dyq;
                            lz = s * dzq;
                        } else if (mode == 3) {
                            float s = -corr;
                            lx = s * dxq;
                            ly = s * dyq;
                            lz = s * dzq;
                        } else if (mode == 4) {
                            float s = -(value - reflected - corr);
                            lx = s * dxq;
                            ly = s * dyq;
                            lz = s * dzq;
                        } else if (mode == 5) {
                            float s = value - reflected;
                            lx = s * wvx;
                            ly = s * wvy;
                            lz = s * wvz;
                        }
                        flx = flx + lx;
                        fly = fly + ly;
                        flz = flz + lz;
                        const float bpx = wx - lf * dxq * p.dx - bx;
                        const float bpy = wy - lf * dyq * p.dx - by;
                        const float bpz = wz - lf * dzq * p.dx - bz;
                        float ox, oy, oz;
                        cross3(bpx, bpy, bpz, lx * fscale, ly * fscale, lz * fscale, ox, oy, oz);
                        tqx = tqx + ox;
                        tqy = tqy + oy;
                        tqz = tqz + oz;
                    }
                } else {
                    value = fsrc[q];
                }
            }
            f[q] = value;
            rho = rho + value;
            mx = mx + value * dxq;
            my = my + value * dyq;
            mz = mz + value * dzq;
            });
        };
        link_group(QSeq<0, 1, 2, 3, 4, 5, 6>{}, std::integral_constant<int, 0>{}, 0x7Eu);
        link_group(QSeq<7, 8, 9, 10, 11, 12>{}, std::integral_constant<int, 7>{}, 0x1F80u);
        link_group(QSeq<13, 14, 15, 16, 17, 18>{}, std::integral_constant<int, 13>{}, 0x7E000u);
    }

    const float sponge = p.sponge_strength > 0.0f ? sponge_factor(p, x, y, z) : 0.0f;
    float post[19];
    float act, restr;
    if constexpr (ARITH == ARITH_FUSED) {
        collide_fused<COLL>(f, p.omega_plus, p.omega_minus, post, act, restr);
    } else {
        rho = fmaxf(rho, 1.0e-8f);
        const float ux = mx / rho, uy = my / rho, uz = mz / rho;
        collide<COLL>(f, rho, ux, uy, uz, p.omega_plus, p.omega_minus, post, act, restr);
    }
    float crho, cmx, cmy, cmz;
    write_post<S>(fout, x, row, post, sponge, p, crho, cmx, cmy, cmz);
    guard_finite(p, crho);
    if (p.capture_fields) capture(density, velocity, gid, crho, cmx, cmy, cmz);
    if (p.accumulate_loads) {
        L.fx = flx * fscale;
        L.fy = fly * fscale;
        L.fz = flz * fscale;
        L.tx = tqx;
        L.ty = tqy;
        L.tz = tqz;
        L.act = act;
        L.restr = restr;
    }
}

// Block sum of 8 FP64 lanes (sums for 0..6, max for 7) in a fixed tree:
// wavefront shuffles, then 4 wavefront results combined by lane 0.
__device__ __forceinline__ void block_reduce8(double (&v)[8], double* __restrict__ out) {
    __shared__ double wave_part[4][8];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double o = __shfl_down(v[i], off, 64);
            v[i] = i == 7 ? fmax(v[i], o) : v[i] + o;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) wave_part[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double s = wave_part[0][i];
            for (int w = 1; w < (int)(blockDim.x / 64); ++w)
                s = i == 7 ? fmax(s, wave_part[w][i]) : s + wave_part[w][i];
            out[i] = s;
        }
    }
    __syncthreads();
}

template <int ARITH, int COLL, int S>
__global__ __launch_bounds__(256) void fluid_general(const typename Store<S>::T* __restrict__ fin,
                                                     typename Store<S>::T* __restrict__ fout,
                                                     const uint8_t* __restrict__ solid_prev,
                                                     const uint8_t* __restrict__ solid_cur,
                                                     const float4* __restrict__ wall,
                                                     const uint32_t* __restrict__ general_list,
                                                     const uint32_t* __restrict__ general_count,
                                                     double* __restrict__ partials,
                                                     float* __restrict__ density, float4* __restrict__ velocity,
                                                     const double* __restrict__ body, StepParams p) {
    set_priority(p.general_priority);
    const uint32_t count = *general_count;
    // torque reference point: the body position, read once per workgroup
    const float bx = (float)body[0], by = (float)body[1], bz = (float)body[2];
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint32_t tile = general_list[i];
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        CellLoad L{0, 0, 0, 0, 0, 0, 0, 0};
        if (x < p.g.nx && y < p.g.ny)
            general_cell<ARITH, COLL, S>(fin, fout, solid_prev, solid_cur, wall, density, velocity, p, x, y, z, bx, by,
                                             bz, L);
        if (p.accumulate_loads) {
            double v[8] = {(double)L.fx, (double)L.fy, (double)L.fz, (double)L.tx,
                           (double)L.ty, (double)L.tz, (double)L.act, (double)L.restr};
            block_reduce8(v, partials + (size_t)tile * 8);
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from bf_api.hip).
// ---------------------------------------------------------------------------

template <int ARITH, int COLL, int S>
static hipError_t launch_fast_t(const FluidBuffers& b, const StepParams& p, hipStream_t s) {
    using T = typename Store<S>::T;
    if constexpr (S != STORE_FP32) {
        // two cells per lane over tile pairs (whole pairs along x)
        if (p.g.nx % (2 * TILE_X) == 0 && p.pair_cells) {
            hipLaunchKernelGGL((fluid_fast_pair<ARITH, COLL, S>), dim3(p.g.tiles / 2), dim3(256), 0, s,
                               (const T*)b.fin, (T*)b.fout, b.static_flags, b.dyn_prev, b.dyn_cur, b.density,
                               b.velocity, b.activation, p);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((fluid_fast<ARITH, COLL, S>), dim3(p.g.tiles), dim3(256), 0, s, (const T*)b.fin,
                       (T*)b.fout, b.static_flags, b.dyn_prev, b.dyn_cur, b.density, b.velocity, b.activation, p);
    return hipGetLastError();
}

template <int ARITH, int COLL, int S>
static hipError_t launch_general_t(const FluidBuffers& b, const StepParams& p, int general_blocks, hipStream_t s) {
    using T = typename Store<S>::T;
    hipLaunchKernelGGL((fluid_general<ARITH, COLL, S>), dim3(general_blocks), dim3(256), 0, s, (const T*)b.fin,
                       (T*)b.fout, b.solid_prev, b.solid_cur, b.wall, b.general_list, b.general_count,
                       b.partials, b.density, b.velocity, b.body, p);
    return hipGetLastError();
}

}  // namespace bf
