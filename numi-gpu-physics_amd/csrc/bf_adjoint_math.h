// Reverse-mode derivatives of the D3Q19 collision and of the moving-wall
// bounce-back value, for the hand-written adjoint of the step.
//
// Host and device, templated on the real type: the FP32 instantiations run
// in the adjoint kernels; the FP64 ones are checked against central finite
// differences on the CPU (bf_adjoint_check.cpp).
//
// Forward operators (the fused pair form of collide_fused in bf_fluid.h,
// which regroups BirdFlow.metal:2888-3174):
//   TRT      post_q = f_q - w+ (f_q^s - feq_q^s) - w- (f_q^a - feq_q^a)
//   REG/RR3  post_q = feq_q + (1 - w+) w_q [4.5 H2:Pi_neq + 13.5 H3.a3],
//            then the positivity limiter post = feq + s (post - feq).
// Pairs (q, q+1) for odd q have c_{q+1} = -c_q; even parts use f_q + f_{q+1},
// odd parts f_q - f_{q+1}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bf_device.h"

namespace bf {
namespace adj {

template <int... Q>
struct Seq {};
template <typename F, int... Q>
__host__ __device__ __forceinline__ void each(Seq<Q...>, F&& fn) {
    (fn(std::integral_constant<int, Q>{}), ...);
}
using Pairs = Seq<0, 1, 2, 3, 4, 5, 6, 7, 8>;   // pair k: q = 2k + 1 and q + 1
using AllQ = Seq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18>;

template <typename R>
__host__ __device__ __forceinline__ R fma_(R a, R b, R c) {
    if constexpr (std::is_same_v<R, float>) return fmaf(a, b, c);
    else return fma(a, b, c);
}
template <typename R>
__host__ __device__ __forceinline__ R max_(R a, R b) { return a > b ? a : b; }
template <typename R>
__host__ __device__ __forceinline__ R min_(R a, R b) { return a < b ? a : b; }
template <typename R>
__host__ __device__ __forceinline__ R abs_(R a) { return a < R(0) ? -a : a; }

template <int q, typename R>
__host__ __device__ __forceinline__ R cdot(R ux, R uy, R uz) {
    R r = R(0);
    if constexpr (cx(q) != 0) r = r + (cx(q) > 0 ? ux : -ux);
    if constexpr (cy(q) != 0) r = r + (cy(q) > 0 ? uy : -uy);
    if constexpr (cz(q) != 0) r = r + (cz(q) > 0 ? uz : -uz);
    return r;
}

template <typename R>
__host__ __device__ __forceinline__ R wq(int q) {
    return q == 0 ? R(1) / R(3) : (q <= 6 ? R(1) / R(18) : R(1) / R(36));
}

// Intermediates of the fused collision shared by the forward twin and the
// reverse pass.
template <typename R>
struct Fwd {
    R sp[9], dp[9], A[9], B[9], cu[9];
    R rho, inv, ux, uy, uz, base, feq0;
    bool rho_clamped;
    // regularized / RR3
    R nk[9], S, Px, Py, Pz, od0, od1, od2, dg0, dg1, dg2, c2T;
    R aXxy, aXxz, aXyy, aXzz, aYyz, aYzz;
    R P[9], Q[9], post0;
    // limiter
    bool limited;
    R scale;
    int qstar;          // direction that set the scale (-1: clamped at 0 or none)
};

template <int COLL, typename R>
__host__ __device__ __forceinline__ void forward(const R (&f)[19], R op, R om, R (&post)[19], Fwd<R>& t) {
    R rho = f[0], mx = R(0), my = R(0), mz = R(0);
    each(Pairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        t.sp[k] = f[q] + f[q + 1];
        t.dp[k] = f[q] - f[q + 1];
        rho = rho + t.sp[k];
        if constexpr (cx(q) != 0) mx = mx + (cx(q) > 0 ? t.dp[k] : -t.dp[k]);
        if constexpr (cy(q) != 0) my = my + (cy(q) > 0 ? t.dp[k] : -t.dp[k]);
        if constexpr (cz(q) != 0) mz = mz + (cz(q) > 0 ? t.dp[k] : -t.dp[k]);
    });
    t.rho_clamped = rho < R(1.0e-8);
    t.rho = max_(rho, R(1.0e-8));
    t.inv = R(1) / t.rho;
    t.ux = mx * t.inv;
    t.uy = my * t.inv;
    t.uz = mz * t.inv;
    t.base = fma_(R(-1.5), fma_(t.ux, t.ux, fma_(t.uy, t.uy, t.uz * t.uz)), R(1));
    each(Pairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        const R wr = t.rho * wq<R>(q);
        t.cu[k] = cdot<q>(t.ux, t.uy, t.uz);
        t.A[k] = wr * fma_(R(4.5) * t.cu[k], t.cu[k], t.base);
        t.B[k] = (R(3) * wr) * t.cu[k];
    });
    t.feq0 = (t.rho * wq<R>(0)) * t.base;
    t.limited = false;
    t.scale = R(1);
    t.qstar = -1;
    if constexpr (COLL == COLL_TRT || COLL == COLL_TRT_SYMMETRIC_LIMITER) {
        post[0] = fma_(-op, f[0] - t.feq0, f[0]);
        R hs[9], ha[9];
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            hs[k] = fma_(R(0.5), t.sp[k], -t.A[k]);
            ha[k] = fma_(R(0.5), t.dp[k], -t.B[k]);
            post[q] = fma_(-om, ha[k], fma_(-op, hs[k], f[q]));
            post[q + 1] = fma_(om, ha[k], fma_(-op, hs[k], f[q + 1]));
        });
        if constexpr (COLL == COLL_TRT_SYMMETRIC_LIMITER) {
            // the symmetric increment si = -w+ (f^s - feq^s) scaled by the
            // largest s <= 1 keeping every population at or above its floor
            // (collide_fused in bf_fluid.h)
            R lo = fma_(R(-1.0e-6), abs_(t.feq0), post[0]);
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                lo = min_(lo, min_(fma_(R(-1.0e-6), abs_(t.A[k] + t.B[k]), post[q]),
                                   fma_(R(-1.0e-6), abs_(t.A[k] - t.B[k]), post[q + 1])));
            });
            if (lo < R(2.0e-12)) {
                R s = R(1);
                int qs = -1;
                auto limit = [&](int q, R fq, R feq, R si, R ai) {
                    if (si < R(0)) {
                        const R b = fq + ai;
                        const R floorv = max_(R(1.0e-12), R(1.0e-6) * max_(feq, R(0)));
                        if (b + si < floorv) {
                            const R adm = (b - floorv) / max_(-si, R(1.0e-30));
                            const R c = min_(max_(adm, R(0)), R(1));
                            if (c < s) {
                                s = c;
                                qs = adm > R(0) && adm < R(1) ? q : -1;
                            }
                        }
                    }
                };
                limit(0, f[0], t.feq0, -op * (f[0] - t.feq0), R(0));
                each(Pairs{}, [&](auto kc) {
                    constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                    limit(q, f[q], t.A[k] + t.B[k], -op * hs[k], -om * ha[k]);
                    limit(q + 1, f[q + 1], t.A[k] - t.B[k], -op * hs[k], om * ha[k]);
                });
                if (s < R(1)) {
                    t.limited = true;
                    t.scale = s;
                    t.qstar = qs;
                    post[0] = fma_(s, -op * (f[0] - t.feq0), f[0]);
                    each(Pairs{}, [&](auto kc) {
                        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                        post[q] = fma_(s, -op * hs[k], f[q]) - om * ha[k];
                        post[q + 1] = fma_(s, -op * hs[k], f[q + 1]) + om * ha[k];
                    });
                }
            }
        }
    } else {
        constexpr bool RR3 = COLL == COLL_RR3;
        const R c2 = R(1) / R(3);
        const R h_one = R(1) - c2;
        t.S = f[0] - t.feq0;
        t.Px = t.Py = t.Pz = R(0);
        t.od0 = t.od1 = t.od2 = R(0);
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            t.nk[k] = fma_(R(-2), t.A[k], t.sp[k]);
            t.S = t.S + t.nk[k];
            if constexpr (cx(q) != 0) t.Px = t.Px + t.nk[k];
            if constexpr (cy(q) != 0) t.Py = t.Py + t.nk[k];
            if constexpr (cz(q) != 0) t.Pz = t.Pz + t.nk[k];
            if constexpr (cx(q) * cy(q) != 0) t.od0 = t.od0 + (cx(q) * cy(q) > 0 ? t.nk[k] : -t.nk[k]);
            if constexpr (cx(q) * cz(q) != 0) t.od1 = t.od1 + (cx(q) * cz(q) > 0 ? t.nk[k] : -t.nk[k]);
            if constexpr (cy(q) * cz(q) != 0) t.od2 = t.od2 + (cy(q) * cz(q) > 0 ? t.nk[k] : -t.nk[k]);
        });
        t.dg0 = fma_(-c2, t.S, t.Px);
        t.dg1 = fma_(-c2, t.S, t.Py);
        t.dg2 = fma_(-c2, t.S, t.Pz);
        t.c2T = c2 * ((t.dg0 + t.dg1) + t.dg2);
        const R om1 = R(1) - op;
        t.aXxy = t.aXxz = t.aXyy = t.aXzz = t.aYyz = t.aYzz = R(0);
        if constexpr (RR3) {
            t.aXxy = fma_(R(2) * t.ux, t.od0, t.uy * t.dg0);
            t.aXxz = fma_(R(2) * t.ux, t.od1, t.uz * t.dg0);
            t.aXyy = fma_(R(2) * t.uy, t.od0, t.ux * t.dg1);
            t.aXzz = fma_(R(2) * t.uz, t.od1, t.ux * t.dg2);
            t.aYyz = fma_(R(2) * t.uy, t.od2, t.uz * t.dg1);
            t.aYzz = fma_(R(2) * t.uz, t.od2, t.uy * t.dg2);
        }
        t.post0 = fma_(R(-4.5) * wq<R>(0) * om1, t.c2T, t.feq0);
        post[0] = t.post0;
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            constexpr int X = cx(q), Y = cy(q), Z = cz(q);
            R contr = -t.c2T;
            if constexpr (X != 0) contr = contr + t.dg0;
            if constexpr (Y != 0) contr = contr + t.dg1;
            if constexpr (Z != 0) contr = contr + t.dg2;
            if constexpr (X * Y != 0) contr = fma_(X * Y > 0 ? R(2) : R(-2), t.od0, contr);
            if constexpr (X * Z != 0) contr = fma_(X * Z > 0 ? R(2) : R(-2), t.od1, contr);
            if constexpr (Y * Z != 0) contr = fma_(Y * Z > 0 ? R(2) : R(-2), t.od2, contr);
            const R w = wq<R>(q);
            t.P[k] = fma_((R(4.5) * w) * om1, contr, t.A[k]);
            R Qv = t.B[k];
            if constexpr (RR3) {
                // H3: sum of c_lead h_other a over the nonzero leads;
                // h = 1 - cs2 on a nonzero axis, -cs2 on a zero one
                R t_one = R(0), t_zero = R(0);
                auto term = [&](int lead, bool other_nonzero, R a) {
                    if (lead == 0) return;
                    const R v = lead > 0 ? a : -a;
                    if (other_nonzero) t_one = t_one + v;
                    else t_zero = t_zero + v;
                };
                term(Y, X != 0, t.aXxy);
                term(Z, X != 0, t.aXxz);
                term(X, Y != 0, t.aXyy);
                term(X, Z != 0, t.aXzz);
                term(Z, Y != 0, t.aYyz);
                term(Y, Z != 0, t.aYzz);
                const R third = fma_(h_one, t_one, -c2 * t_zero);
                Qv = fma_((R(13.5) * w) * om1, third, Qv);
            }
            t.Q[k] = Qv;
            post[q] = t.P[k] + Qv;
            post[q + 1] = t.P[k] - Qv;
        });
        // positivity limiter
        R lo = fma_(R(-1.0e-6), abs_(t.feq0), post[0]);
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            lo = min_(lo, min_(fma_(R(-1.0e-6), abs_(t.A[k] + t.B[k]), post[q]),
                               fma_(R(-1.0e-6), abs_(t.A[k] - t.B[k]), post[q + 1])));
        });
        if (lo < R(2.0e-12)) {
            R s = R(1);
            int qs = -1;
            auto limit = [&](int q, R feq, R cand) {
                const R floorv = max_(R(1.0e-12), R(1.0e-6) * max_(feq, R(0)));
                if (cand < floorv) {
                    const R adm = (feq - floorv) / max_(feq - cand, R(1.0e-30));
                    const R c = min_(max_(adm, R(0)), R(1));
                    if (c < s) {
                        s = c;
                        qs = adm > R(0) && adm < R(1) ? q : -1;
                    }
                }
            };
            limit(0, t.feq0, post[0]);
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                limit(q, t.A[k] + t.B[k], post[q]);
                limit(q + 1, t.A[k] - t.B[k], post[q + 1]);
            });
            if (s < R(1)) {
                t.limited = true;
                t.scale = s;
                t.qstar = qs;
                post[0] = fma_(s, post[0] - t.feq0, t.feq0);
                each(Pairs{}, [&](auto kc) {
                    constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                    post[q] = fma_(s, post[q] - (t.A[k] + t.B[k]), t.A[k] + t.B[k]);
                    post[q + 1] = fma_(s, post[q + 1] - (t.A[k] - t.B[k]), t.A[k] - t.B[k]);
                });
            }
        }
    }
}

// lam = (d post / d f)^T lp at f.  TRT, regularized and RR3 (with the
// positivity limiter, differentiated through the direction that sets its
// scale).
template <int COLL, typename R>
__host__ __device__ __forceinline__ void reverse(const R (&f)[19], R op, R om, const R (&lp)[19], R (&lam)[19]) {
    Fwd<R> t;
    R post[19];
    forward<COLL>(f, op, om, post, t);
    R lA[9], lB[9], lsp[9], ldp[9];
    R lfeq0 = R(0), lf0 = R(0);
    R direct[19];
    each(AllQ{}, [&](auto qc) { direct[decltype(qc)::value] = R(0); });
    each(Pairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        lA[k] = lB[k] = lsp[k] = ldp[k] = R(0);
    });
    if constexpr (COLL == COLL_TRT) {
        lf0 = (R(1) - op) * lp[0];
        lfeq0 = op * lp[0];
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            direct[q] = lp[q];
            direct[q + 1] = lp[q + 1];
            const R lh = -op * (lp[q] + lp[q + 1]);
            const R lhd = -om * (lp[q] - lp[q + 1]);
            lsp[k] = R(0.5) * lh;
            lA[k] = -lh;
            ldp[k] = R(0.5) * lhd;
            lB[k] = -lhd;
        });
    } else if constexpr (COLL == COLL_TRT_SYMMETRIC_LIMITER) {
        // post_q = f_q + s si_q + ai_q with si = -w+ h (h: symmetric
        // non-equilibrium), ai = -w- hd; s = 1 unless limited, then
        // s = (f* + ai* - floor*) / (-si*) of the direction q* that sets it
        const R s = t.scale;
        R lh[9], lhd[9];
        R lsi0 = s * lp[0];
        R ls = R(0);
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            lh[k] = -op * s * (lp[q] + lp[q + 1]);
            lhd[k] = -om * (lp[q] - lp[q + 1]);
        });
        lf0 = lp[0];
        if (t.limited) {
            const R si0 = -op * (f[0] - t.feq0);
            ls = si0 * lp[0];
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                ls = ls + (-op * fma_(R(0.5), t.sp[k], -t.A[k])) * (lp[q] + lp[q + 1]);
            });
        }
        R lfeq_star = R(0);
        if (t.limited && t.qstar >= 0) {
            // the binding direction's population, increments and equilibrium,
            // gathered by compile-time index (a run-time index into these
            // register arrays would place them in scratch memory)
            const int qs = t.qstar;
            R fq = f[0], feq = t.feq0, si = -op * (f[0] - t.feq0), ai = R(0);
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                const R hsk = fma_(R(0.5), t.sp[k], -t.A[k]);
                const R hak = fma_(R(0.5), t.dp[k], -t.B[k]);
                if (qs == q) {              // +B, -w- hd
                    fq = f[q];
                    feq = t.A[k] + t.B[k];
                    si = -op * hsk;
                    ai = -om * hak;
                } else if (qs == q + 1) {   // -B, +w- hd
                    fq = f[q + 1];
                    feq = t.A[k] - t.B[k];
                    si = -op * hsk;
                    ai = om * hak;
                }
            });
            const bool lin = R(1.0e-6) * max_(feq, R(0)) > R(1.0e-12);
            const R fl = lin ? R(1.0e-6) * feq : R(1.0e-12);
            const R num = fq + ai - fl;
            const R ds_dnum = R(1) / (-si);
            const R ds_dsi = num / (si * si);
            const R lnum = ls * ds_dnum;
            const R lsi = ls * ds_dsi;
            lfeq_star = lin ? -R(1.0e-6) * lnum : R(0);
            // num = f* + ai* - floor(feq*); si* = -w+ (f0 - feq0) or -w+ hs_k
            if (qs == 0) {
                lf0 = lf0 + lnum;
                lf0 = lf0 - op * lsi;
                lfeq0 = lfeq0 + op * lsi;
                lfeq0 = lfeq0 + lfeq_star;
            }
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                if (qs == q || qs == q + 1) {
                    const R sgn = qs == q ? R(1) : R(-1);
                    if (qs == q) direct[q] = direct[q] + lnum;
                    else direct[q + 1] = direct[q + 1] + lnum;
                    lhd[k] = lhd[k] + (-sgn * om) * lnum;
                    lh[k] = lh[k] - op * lsi;
                    lA[k] = lA[k] + lfeq_star;
                    lB[k] = lB[k] + sgn * lfeq_star;
                }
            });
        }
        // post_0 = f0 + s (-w+)(f0 - feq0)
        lf0 = lf0 - op * lsi0;
        lfeq0 = lfeq0 + op * lsi0;
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            direct[q] = direct[q] + lp[q];
            direct[q + 1] = direct[q + 1] + lp[q + 1];
            lsp[k] = R(0.5) * lh[k];
            lA[k] = lA[k] - lh[k];
            ldp[k] = R(0.5) * lhd[k];
            lB[k] = lB[k] - lhd[k];
        });
    } else {
        constexpr bool RR3 = COLL == COLL_RR3;
        const R c2 = R(1) / R(3);
        const R h_one = R(1) - c2;
        const R om1 = R(1) - op;
        // limiter: post = feq + s (cand - feq)
        R lcand[19], lfeq[19];
        each(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lcand[q] = lp[q];
            lfeq[q] = R(0);
        });
        if (t.limited) {
            R cand[19], feq[19];
            cand[0] = t.post0;
            feq[0] = t.feq0;
            each(Pairs{}, [&](auto kc) {
                constexpr int k = decltype(kc)::value, q = 2 * k + 1;
                cand[q] = t.P[k] + t.Q[k];
                cand[q + 1] = t.P[k] - t.Q[k];
                feq[q] = t.A[k] + t.B[k];
                feq[q + 1] = t.A[k] - t.B[k];
            });
            R ls = R(0);
            each(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                ls = ls + (cand[q] - feq[q]) * lp[q];
                lcand[q] = t.scale * lp[q];
                lfeq[q] = (R(1) - t.scale) * lp[q];
            });
            if (t.qstar >= 0) {
                each(AllQ{}, [&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if (q == t.qstar) {
                        const R fe = feq[q], ca = cand[q];
                        const bool lin = R(1.0e-6) * fe > R(1.0e-12);
                        const R fl = lin ? R(1.0e-6) * fe : R(1.0e-12);
                        const R d = fe - ca, n = fe - fl;
                        lfeq[q] = lfeq[q] + ls * ((R(1) - (lin ? R(1.0e-6) : R(0))) * d - n) / (d * d);
                        lcand[q] = lcand[q] + ls * n / (d * d);
                    }
                });
            }
        }
        // cand0 = feq0 - 4.5 w0 om1 c2T ; cand_q, cand_{q+1} = P +- Q
        R lc2T = R(-4.5) * wq<R>(0) * om1 * lcand[0];
        lfeq0 = lcand[0] + lfeq[0];
        R ldg0 = R(0), ldg1 = R(0), ldg2 = R(0), lod0 = R(0), lod1 = R(0), lod2 = R(0);
        R laXxy = R(0), laXxz = R(0), laXyy = R(0), laXzz = R(0), laYyz = R(0), laYzz = R(0);
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            constexpr int X = cx(q), Y = cy(q), Z = cz(q);
            const R w = wq<R>(q);
            const R lP = lcand[q] + lcand[q + 1];
            const R lQ = lcand[q] - lcand[q + 1];
            lA[k] = lP + (lfeq[q] + lfeq[q + 1]);
            lB[k] = lQ + (lfeq[q] - lfeq[q + 1]);
            const R lcontr = (R(4.5) * w) * om1 * lP;
            lc2T = lc2T - lcontr;
            if constexpr (X != 0) ldg0 = ldg0 + lcontr;
            if constexpr (Y != 0) ldg1 = ldg1 + lcontr;
            if constexpr (Z != 0) ldg2 = ldg2 + lcontr;
            if constexpr (X * Y != 0) lod0 = lod0 + (X * Y > 0 ? R(2) : R(-2)) * lcontr;
            if constexpr (X * Z != 0) lod1 = lod1 + (X * Z > 0 ? R(2) : R(-2)) * lcontr;
            if constexpr (Y * Z != 0) lod2 = lod2 + (Y * Z > 0 ? R(2) : R(-2)) * lcontr;
            if constexpr (RR3) {
                const R lthird = (R(13.5) * w) * om1 * lQ;
                auto term = [&](int lead, bool other_nonzero, R& la) {
                    if (lead == 0) return;
                    const R g = (other_nonzero ? h_one : -c2) * lthird;
                    la = la + (lead > 0 ? g : -g);
                };
                term(Y, X != 0, laXxy);
                term(Z, X != 0, laXxz);
                term(X, Y != 0, laXyy);
                term(X, Z != 0, laXzz);
                term(Z, Y != 0, laYyz);
                term(Y, Z != 0, laYzz);
            }
        });
        R lux = R(0), luy = R(0), luz = R(0);
        if constexpr (RR3) {
            // aXxy = 2ux od0 + uy dg0, ...
            lux = lux + R(2) * t.od0 * laXxy + R(2) * t.od1 * laXxz + t.dg1 * laXyy + t.dg2 * laXzz;
            luy = luy + t.dg0 * laXxy + R(2) * t.od0 * laXyy + R(2) * t.od2 * laYyz + t.dg2 * laYzz;
            luz = luz + t.dg0 * laXxz + R(2) * t.od1 * laXzz + t.dg1 * laYyz + R(2) * t.od2 * laYzz;
            lod0 = lod0 + R(2) * t.ux * laXxy + R(2) * t.uy * laXyy;
            lod1 = lod1 + R(2) * t.ux * laXxz + R(2) * t.uz * laXzz;
            lod2 = lod2 + R(2) * t.uy * laYyz + R(2) * t.uz * laYzz;
            ldg0 = ldg0 + t.uy * laXxy + t.uz * laXxz;
            ldg1 = ldg1 + t.ux * laXyy + t.uz * laYyz;
            ldg2 = ldg2 + t.ux * laXzz + t.uy * laYzz;
        }
        // c2T = c2 (dg0 + dg1 + dg2); dg_a = P_a - c2 S
        ldg0 = ldg0 + c2 * lc2T;
        ldg1 = ldg1 + c2 * lc2T;
        ldg2 = ldg2 + c2 * lc2T;
        const R lS = -c2 * (ldg0 + ldg1 + ldg2);
        // S = (f0 - feq0) + sum nk; P_a, od from nk; nk = sp - 2A
        lf0 = lS;
        lfeq0 = lfeq0 - lS;
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            R lnk = lS;
            if constexpr (cx(q) != 0) lnk = lnk + ldg0;
            if constexpr (cy(q) != 0) lnk = lnk + ldg1;
            if constexpr (cz(q) != 0) lnk = lnk + ldg2;
            if constexpr (cx(q) * cy(q) != 0) lnk = lnk + (cx(q) * cy(q) > 0 ? lod0 : -lod0);
            if constexpr (cx(q) * cz(q) != 0) lnk = lnk + (cx(q) * cz(q) > 0 ? lod1 : -lod1);
            if constexpr (cy(q) * cz(q) != 0) lnk = lnk + (cy(q) * cz(q) > 0 ? lod2 : -lod2);
            lsp[k] = lnk;
            lA[k] = lA[k] - R(2) * lnk;
        });
        // velocity adjoint from the third-order coefficients joins below
        // through lux, luy, luz
        R lrho = R(0), lbase = R(0);
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            const R w = wq<R>(q), wr = t.rho * w, cu = t.cu[k];
            lrho = lrho + w * ((R(4.5) * cu * cu + t.base) * lA[k] + R(3) * cu * lB[k]);
            const R lcu = wr * (R(9) * cu * lA[k] + R(3) * lB[k]);
            lbase = lbase + wr * lA[k];
            if constexpr (cx(q) != 0) lux = lux + (cx(q) > 0 ? lcu : -lcu);
            if constexpr (cy(q) != 0) luy = luy + (cy(q) > 0 ? lcu : -lcu);
            if constexpr (cz(q) != 0) luz = luz + (cz(q) > 0 ? lcu : -lcu);
        });
        lrho = lrho + wq<R>(0) * t.base * lfeq0;
        lbase = lbase + t.rho * wq<R>(0) * lfeq0;
        const R lusq = R(-1.5) * lbase;
        lux = lux + R(2) * t.ux * lusq;
        luy = luy + R(2) * t.uy * lusq;
        luz = luz + R(2) * t.uz * lusq;
        lrho = lrho - (lux * t.ux + luy * t.uy + luz * t.uz) * t.inv;
        const R lmx = lux * t.inv, lmy = luy * t.inv, lmz = luz * t.inv;
        if (t.rho_clamped) lrho = R(0);
        lf0 = lf0 + lrho;
        each(Pairs{}, [&](auto kc) {
            constexpr int k = decltype(kc)::value, q = 2 * k + 1;
            R ld = R(0);
            if constexpr (cx(q) != 0) ld = ld + (cx(q) > 0 ? lmx : -lmx);
            if constexpr (cy(q) != 0) ld = ld + (cy(q) > 0 ? lmy : -lmy);
            if constexpr (cz(q) != 0) ld = ld + (cz(q) > 0 ? lmz : -lmz);
            const R ls_ = lsp[k] + lrho;
            lam[q] = ls_ + ld;
            lam[q + 1] = ls_ - ld;
        });
        lam[0] = lf0;
        return;
    }
    // TRT: equilibrium reverse
    R lux = R(0), luy = R(0), luz = R(0), lrho = R(0), lbase = R(0);
    each(Pairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        const R w = wq<R>(q), wr = t.rho * w, cu = t.cu[k];
        lrho = lrho + w * ((R(4.5) * cu * cu + t.base) * lA[k] + R(3) * cu * lB[k]);
        const R lcu = wr * (R(9) * cu * lA[k] + R(3) * lB[k]);
        lbase = lbase + wr * lA[k];
        if constexpr (cx(q) != 0) lux = lux + (cx(q) > 0 ? lcu : -lcu);
        if constexpr (cy(q) != 0) luy = luy + (cy(q) > 0 ? lcu : -lcu);
        if constexpr (cz(q) != 0) luz = luz + (cz(q) > 0 ? lcu : -lcu);
    });
    lrho = lrho + wq<R>(0) * t.base * lfeq0;
    lbase = lbase + t.rho * wq<R>(0) * lfeq0;
    const R lusq = R(-1.5) * lbase;
    lux = lux + R(2) * t.ux * lusq;
    luy = luy + R(2) * t.uy * lusq;
    luz = luz + R(2) * t.uz * lusq;
    lrho = lrho - (lux * t.ux + luy * t.uy + luz * t.uz) * t.inv;
    const R lmx = lux * t.inv, lmy = luy * t.inv, lmz = luz * t.inv;
    if (t.rho_clamped) lrho = R(0);
    lf0 = lf0 + lrho;
    each(Pairs{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value, q = 2 * k + 1;
        R ld = ldp[k];
        if constexpr (cx(q) != 0) ld = ld + (cx(q) > 0 ? lmx : -lmx);
        if constexpr (cy(q) != 0) ld = ld + (cy(q) > 0 ? lmy : -lmy);
        if constexpr (cz(q) != 0) ld = ld + (cz(q) > 0 ? lmz : -lmz);
        const R ls_ = lsp[k] + lrho;
        lam[q] = direct[q] + ls_ + ld;
        lam[q + 1] = direct[q + 1] + ls_ - ld;
    });
    lam[0] = lf0;
}

// d feq_q(1, u) / du = w_q (3 c_q + 9 (c_q.u) c_q - 3 u), contracted with
// the adjoint of the populations: g = sum_q l_q d feq_q / du.
template <typename R>
__host__ __device__ __forceinline__ void feq_velocity_adjoint(const R (&l)[19], R ux, R uy, R uz, R& gx, R& gy,
                                                             R& gz) {
    gx = gy = gz = R(0);
    each(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const R w = wq<R>(q), cu = cdot<q>(ux, uy, uz);
        const R s = w * l[q];
        gx = gx + s * (R(3) * R(cx(q)) + R(9) * cu * R(cx(q)) - R(3) * ux);
        gy = gy + s * (R(3) * R(cy(q)) + R(9) * cu * R(cy(q)) - R(3) * uy);
        gz = gz + s * (R(3) * R(cz(q)) + R(9) * cu * R(cz(q)) - R(3) * uz);
    });
}

// Moving-wall bounce-back value of one link (BirdFlow.metal:2729-2806) and
// its partial derivatives.  Inputs: reflected = f_opp(x), farther =
// f_opp(x + c_q), pin = f_q(x); wg / ws = wall entries (velocity xyz,
// signed distance w) of the fluid cell and of the solid source; rule:
// WALL_*; farther_ok: x + c_q is in the domain and fluid.
template <typename R>
struct LinkD {
    R value;
    R d_refl, d_farther, d_pin;
    R d_wg[4], d_ws[4];   // d value / d (wall xyz, w) of the fluid cell and of the source
    R lf;
    R dv_dlf;             // d value / d lf (the link-table rule differentiates lf itself)
    bool uses_farther, uses_pin;
};

// table_lf: the link's fraction under WALL_INTERPOLATED_LINKS (clamped like
// the step); ignored by the other rules.
template <typename R>
__host__ __device__ __forceinline__ LinkD<R> link_value(int q_cx, int q_cy, int q_cz, R w_q, R rho_wall, int rule,
                                                        bool farther_ok, R reflected, R farther, R pin,
                                                        const R (&wg)[4], const R (&ws)[4], R table_lf = R(0)) {
    LinkD<R> o;
    for (int i = 0; i < 4; ++i) o.d_wg[i] = o.d_ws[i] = R(0);
    o.d_refl = o.d_farther = o.d_pin = R(0);
    o.uses_farther = o.uses_pin = false;
    R lf = R(0.5), dlf_dwgw = R(0), dlf_dwsw = R(0);
    bool interp = rule != WALL_HALFWAY;
    if (interp && rule == WALL_INTERPOLATED_SDF) {
        const R df = max_(wg[3], R(0)), ds = min_(ws[3], R(0));
        const R den = df - ds;
        const R raw = df / max_(den, R(1.0e-6));
        lf = min_(max_(raw, R(1.0e-4)), R(1));
        if (den > R(1.0e-6) && raw > R(1.0e-4) && raw < R(1)) {
            const R inv2 = R(1) / (den * den);
            dlf_dwgw = wg[3] > R(0) ? -ds * inv2 : R(0);
            dlf_dwsw = ws[3] < R(0) ? df * inv2 : R(0);
        }
    } else if (interp && rule == WALL_INTERPOLATED_LINKS) {
        lf = min_(max_(table_lf, R(1.0e-4)), R(1));
    }
    if (interp && lf <= R(0.5) && !farther_ok) {
        interp = false;   // halfway fallback
        lf = R(0.5);
        dlf_dwgw = dlf_dwsw = R(0);
    }
    R wv[3], dwv_dlf[3];
    for (int i = 0; i < 3; ++i) {
        wv[i] = interp ? wg[i] + (ws[i] - wg[i]) * lf : ws[i];
        dwv_dlf[i] = interp ? ws[i] - wg[i] : R(0);
    }
    const R kc = R(6) * w_q * rho_wall;   // 2 w rho_w / cs2
    const R cq[3] = {R(q_cx), R(q_cy), R(q_cz)};
    const R corr = kc * (cq[0] * wv[0] + cq[1] * wv[1] + cq[2] * wv[2]);
    const R dcorr_dlf = kc * (cq[0] * dwv_dlf[0] + cq[1] * dwv_dlf[1] + cq[2] * dwv_dlf[2]);
    R dv_dcorr = R(1), dv_dlf = R(0);
    if (!interp) {
        o.value = reflected + corr;
        o.d_refl = R(1);
    } else if (lf <= R(0.5)) {
        o.value = R(2) * lf * reflected + (R(1) - R(2) * lf) * farther + corr;
        o.d_refl = R(2) * lf;
        o.d_farther = R(1) - R(2) * lf;
        o.uses_farther = true;
        dv_dlf = R(2) * reflected - R(2) * farther + dcorr_dlf;
    } else {
        const R i2 = R(1) / (R(2) * lf);
        o.value = (reflected + corr) * i2 + (R(2) * lf - R(1)) * pin * i2;
        o.d_refl = i2;
        o.d_pin = (R(2) * lf - R(1)) * i2;
        o.uses_pin = true;
        dv_dcorr = i2;
        dv_dlf = -(reflected + corr) * i2 / lf + dcorr_dlf * i2 + pin * i2 / lf;
    }
    // wall velocity: wv = wg + (ws - wg) lf (interpolated) or ws (halfway)
    for (int i = 0; i < 3; ++i) {
        const R dcorr_dwv = kc * cq[i];
        o.d_wg[i] = dv_dcorr * dcorr_dwv * (interp ? R(1) - lf : R(0));
        o.d_ws[i] = dv_dcorr * dcorr_dwv * (interp ? lf : R(1));
    }
    o.d_wg[3] = dv_dlf * dlf_dwgw;
    o.d_ws[3] = dv_dlf * dlf_dwsw;
    o.lf = lf;
    o.dv_dlf = dv_dlf;
    return o;
}

}  // namespace adj
}  // namespace bf
