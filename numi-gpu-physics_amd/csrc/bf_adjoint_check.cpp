// FP64 host instantiations of the adjoint math (bf_adjoint_math.h) behind a
// C ABI, so that tests can check every reverse derivative against central
// finite differences of its forward twin.  Test support; not linked into
// libbirdflow_hip.so.
#include "bf_adjoint_math.h"

using namespace bf;

namespace {
template <int C>
void fwd(const double* f, double op, double om, double* post) {
    double fi[19], po[19];
    for (int q = 0; q < 19; ++q) fi[q] = f[q];
    adj::Fwd<double> t;
    adj::forward<C>(fi, op, om, po, t);
    for (int q = 0; q < 19; ++q) post[q] = po[q];
}
template <int C>
void rev(const double* f, double op, double om, const double* lp, double* lam) {
    double fi[19], l[19], o[19];
    for (int q = 0; q < 19; ++q) {
        fi[q] = f[q];
        l[q] = lp[q];
    }
    adj::reverse<C>(fi, op, om, l, o);
    for (int q = 0; q < 19; ++q) lam[q] = o[q];
}
}  // namespace

extern "C" {

int bfa_collide_f64(int coll, const double* f, double op, double om, double* post, int* limited) {
    double fi[19], po[19];
    for (int q = 0; q < 19; ++q) fi[q] = f[q];
    adj::Fwd<double> t;
    switch (coll) {
        case COLL_TRT: adj::forward<COLL_TRT>(fi, op, om, po, t); break;
        case COLL_REGULARIZED: adj::forward<COLL_REGULARIZED>(fi, op, om, po, t); break;
        case COLL_RR3: adj::forward<COLL_RR3>(fi, op, om, po, t); break;
        default: return 1;
    }
    for (int q = 0; q < 19; ++q) post[q] = po[q];
    if (limited) *limited = t.limited ? 1 : 0;
    return 0;
}

int bfa_collide_adjoint_f64(int coll, const double* f, double op, double om, const double* lp, double* lam) {
    switch (coll) {
        case COLL_TRT: rev<COLL_TRT>(f, op, om, lp, lam); return 0;
        case COLL_REGULARIZED: rev<COLL_REGULARIZED>(f, op, om, lp, lam); return 0;
        case COLL_RR3: rev<COLL_RR3>(f, op, om, lp, lam); return 0;
        default: return 1;
    }
}

// out: value, d_refl, d_farther, d_pin, d_wg[4], d_ws[4], lf
void bfa_link_f64(int q, int rule, int farther_ok, double rho_wall, const double* refl_farther_pin, const double* wg,
                  const double* ws, double* out) {
    const double g[4] = {wg[0], wg[1], wg[2], wg[3]}, s[4] = {ws[0], ws[1], ws[2], ws[3]};
    const adj::LinkD<double> d = adj::link_value<double>(cx(q), cy(q), cz(q), adj::wq<double>(q), rho_wall, rule,
                                                         farther_ok != 0, refl_farther_pin[0], refl_farther_pin[1],
                                                         refl_farther_pin[2], g, s);
    out[0] = d.value;
    out[1] = d.d_refl;
    out[2] = d.d_farther;
    out[3] = d.d_pin;
    for (int i = 0; i < 4; ++i) {
        out[4 + i] = d.d_wg[i];
        out[8 + i] = d.d_ws[i];
    }
    out[12] = d.lf;
}

void bfa_feq_velocity_adjoint_f64(const double* l, const double* u, double* g) {
    double li[19];
    for (int q = 0; q < 19; ++q) li[q] = l[q];
    adj::feq_velocity_adjoint<double>(li, u[0], u[1], u[2], g[0], g[1], g[2]);
}

double bfa_equilibrium_f64(int q, double rho, double ux, double uy, double uz) {
    return equilibrium<double>(q, rho, ux, uy, uz);
}

}  // extern "C"
