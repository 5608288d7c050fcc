// FP64 host instantiations of the adjoint math (bf_adjoint_math.h) behind a
// C ABI, so that tests can check every reverse derivative against central
// finite differences of its forward twin, and the host instantiation of the
// population-layout index functions, so that tests can check every index
// they produce against its allocation without a GPU.  Test support; not
// linked into libbirdflow_hip.so.
#include <algorithm>
#include <vector>

#include "bf_adjoint_math.h"
#include "bf_fluid.h"
#include "bf_rays.h"

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
        case COLL_TRT_SYMMETRIC_LIMITER: adj::forward<COLL_TRT_SYMMETRIC_LIMITER>(fi, op, om, po, t); break;
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
        case COLL_TRT_SYMMETRIC_LIMITER: rev<COLL_TRT_SYMMETRIC_LIMITER>(f, op, om, lp, lam); return 0;
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

// The link-table rule (lf from the table): out = value, d value / d lf, lf
void bfa_link_table_f64(int q, int farther_ok, double rho_wall, const double* refl_farther_pin, const double* wg,
                        const double* ws, double table_lf, double* out) {
    const double g[4] = {wg[0], wg[1], wg[2], wg[3]}, s[4] = {ws[0], ws[1], ws[2], ws[3]};
    const adj::LinkD<double> d = adj::link_value<double>(
        cx(q), cy(q), cz(q), adj::wq<double>(q), rho_wall, WALL_INTERPOLATED_LINKS, farther_ok != 0,
        refl_farther_pin[0], refl_farther_pin[1], refl_farther_pin[2], g, s, table_lf);
    out[0] = d.value;
    out[1] = d.dv_dlf;
    out[2] = d.lf;
}

void bfa_feq_velocity_adjoint_f64(const double* l, const double* u, double* g) {
    double li[19];
    for (int q = 0; q < 19; ++q) li[q] = l[q];
    adj::feq_velocity_adjoint<double>(li, u[0], u[1], u[2], g[0], g[1], g[2]);
}

double bfa_equilibrium_f64(int q, double rho, double ux, double uy, double uz) {
    return equilibrium<double>(q, rho, ux, uy, uz);
}

// Index bounds of the population layout on an nx x ny x nz grid with face
// types bc (3 x 2, BC_*), through the same index functions the kernels use
// (pop_index, tile_origin, xcd_remap, shift_axis), along each kernel's own
// index path:
//   pack / unpack / permute: i -> (q, cell, row, x) -> pop_index, for i at
//     both ends of every direction plane and at a stride through the rest;
//   fluid_fast / fluid_general: block -> tile (xcd_remap) -> cell (lane) ->
//     the 19 pull sources (face treatment) -> pop_index, for every lane of
//     the tiles at the ends of every tile row, plane and of the grid;
//   fluid_fast_pair (2-byte storage): block -> tile pair -> row -> the
//     aligned 4-byte words of the lanes, the halo words and the stores;
//   the wing link table: (tile, q, lane) -> tile * 19 * 256 + q * 256 + lane;
//   xcd_remap: a bijection of [0, tiles) (checked exhaustively).
// out[0] = largest population index produced, out[1] = population elements
// allocated (19 N), out[2] = largest link-table index, out[3] = its
// elements, out[4] = pull sources checked.  Returns 0 when every index is
// inside its allocation and the remap is a bijection; otherwise the number of
// the first failing check (1 pack path, 2 fluid path, 3 link table, 4 remap,
// 5 cell index beyond 32 bits, 6 the two-cells-per-lane path: a word beyond
// the lattice or a row not starting at an even element).
int bfa_layout_check(uint32_t nx, uint32_t ny, uint32_t nz, const int* bc, uint64_t* out) {
    Grid g{};
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.nz = (int)nz;
    g.tiles_x = (g.nx + TILE_X - 1) / TILE_X;
    g.tiles_y = (g.ny + TILE_Y - 1) / TILE_Y;
    const uint64_t n64 = (uint64_t)nx * ny * nz;
    if (n64 >= (1ull << 32)) return 5;
    g.n = (uint32_t)n64;
    g.tiles = (uint32_t)((uint64_t)g.tiles_x * g.tiles_y * g.nz);
    const uint64_t elements = 19ull * n64, table = (uint64_t)g.tiles * 19u * TILE_CELLS;
    uint64_t max_pop = 0, max_link = 0, sources = 0;
    int fail = 0;
    auto pop = [&](size_t idx, int path) {
        if (idx > max_pop) max_pop = idx;
        if (idx >= elements && !fail) fail = path;
    };
    // pack / unpack / permute
    const uint64_t stride = std::max<uint64_t>(1, elements / 65536);
    auto pack_at = [&](uint64_t i) {
        if (i >= elements) return;
        const int q = (int)(i / g.n);
        const uint32_t gid = (uint32_t)(i - (uint64_t)q * g.n);
        const uint32_t row = gid / (uint32_t)g.nx;
        pop(pop_index<STORE_FP32>(g, q, (int)(gid - row * (uint32_t)g.nx), row), 1);
    };
    for (int q = 0; q < 19; ++q)
        for (uint64_t k = 0; k < 4096; ++k) {
            pack_at((uint64_t)q * n64 + k);
            pack_at((uint64_t)(q + 1) * n64 - 1 - k);
        }
    for (uint64_t i = 0; i < elements; i += stride) pack_at(i);
    // fluid kernels
    std::vector<uint32_t> tiles;
    const uint32_t per_plane = (uint32_t)g.tiles_x * (uint32_t)g.tiles_y;
    for (uint32_t zc : {0u, 1u, nz / 2, nz - 1})
        for (uint32_t ty : {0u, (uint32_t)g.tiles_y - 1})
            for (uint32_t tx : {0u, (uint32_t)g.tiles_x - 1})
                if (zc < nz) tiles.push_back(zc * per_plane + ty * (uint32_t)g.tiles_x + tx);
    tiles.push_back(xcd_remap(0, g.tiles));
    tiles.push_back(xcd_remap(g.tiles - 1, g.tiles));
    for (uint32_t tile : tiles) {
        int x0, y0, z;
        tile_origin(g, tile, x0, y0, z);
        for (int lane = 0; lane < TILE_CELLS; ++lane) {
            // as fluid_fast: lanes of a partial tile beyond the grid load from
            // the last column / row (the general kernel skips them)
            const int x = std::min(x0 + (lane & (TILE_X - 1)), g.nx - 1), y = std::min(y0 + lane / TILE_X, g.ny - 1);
            for (int q = 0; q < 19; ++q) {
                const uint64_t li = (uint64_t)tile * 19 * TILE_CELLS + (uint64_t)q * TILE_CELLS + lane;
                if (li > max_link) max_link = li;
                if (li >= table && !fail) fail = 3;
            }
            const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
            if ((uint64_t)x + (uint64_t)g.nx * row >= n64 && !fail) fail = 5;
            for (int q = 0; q < 19; ++q) {
                bool far = false;
                const int xs = shift_axis(x, cx(q), g.nx, bc[0], bc[1], far);
                const int ys = shift_axis(y, cy(q), g.ny, bc[2], bc[3], far);
                const int zs = shift_axis(z, cz(q), g.nz, bc[4], bc[5], far);
                pop(pop_index<STORE_FP32>(g, q, x, row), 2);
                ++sources;
                if (far) continue;   // the kernels load the cell itself instead
                pop(pop_index<STORE_FP32>(g, q, xs, (uint32_t)ys + (uint32_t)g.ny * (uint32_t)zs), 2);
            }
        }
    }
    // exact-root pass (bf_rays.h): every puller of a source along q is a
    // grid cell that pulls from that source, with its table slot inside the
    // table; and every cell with a source is among that source's pullers
    {
        const int bcs[3][2] = {{bc[0], bc[1]}, {bc[2], bc[3]}, {bc[4], bc[5]}};
        const int dims[3] = {g.nx, g.ny, g.nz};
        auto cells = [](int n) {
            std::vector<int> v;
            for (int c : {0, 1, n / 2, n - 2, n - 1})
                if (c >= 0 && c < n && std::find(v.begin(), v.end(), c) == v.end()) v.push_back(c);
            return v;
        };
        const std::vector<int> xs = cells(g.nx), ys = cells(g.ny), zs = cells(g.nz);
        for (int x : xs)
            for (int y : ys)
                for (int z : zs)
                    for (int q = 1; q < 19; ++q) {
                        const int c[3] = {cx(q), cy(q), cz(q)};
                        const int at[3] = {x, y, z};
                        int pl[3][2], k[3];
                        for (int a = 0; a < 3; ++a) {
                            k[a] = pullers(bcs[a], at[a], c[a], dims[a], pl[a]);
                            for (int i = 0; i < k[a]; ++i) {
                                int src = -1;
                                if (pl[a][i] < 0 || pl[a][i] >= dims[a] ||
                                    !ray_source(bcs[a], pl[a][i] - c[a], dims[a], src) || src != at[a]) {
                                    if (!fail) fail = 7;
                                }
                            }
                        }
                        for (int iz = 0; iz < k[2]; ++iz)
                            for (int iy = 0; iy < k[1]; ++iy)
                                for (int ix = 0; ix < k[0]; ++ix) {
                                    const size_t li = table_slot(g, pl[0][ix], pl[1][iy], pl[2][iz], q);
                                    if (li > max_link) max_link = li;
                                    if (li >= table && !fail) fail = 7;
                                }
                        // the cell as a puller: its source lists it
                        for (int a = 0; a < 3; ++a) {
                            int src;
                            if (!ray_source(bcs[a], at[a] - c[a], dims[a], src)) continue;
                            int back[2];
                            const int kb = pullers(bcs[a], src, c[a], dims[a], back);
                            bool found = false;
                            for (int i = 0; i < kb; ++i) found = found || back[i] == at[a];
                            if (!found && !fail) fail = 8;
                        }
                    }
    }
    // fluid_fast_pair (2-byte storage, nx a multiple of 128): block -> tile
    // pair -> the 128-cell row of each wavefront -> the aligned 4-byte word
    // of every lane's pair, the halo words at both row ends, the stores
    if (g.nx % (2 * TILE_X) == 0) {
        const uint32_t pairs_x = (uint32_t)g.tiles_x / 2, pairs = g.tiles / 2;
        const uint32_t pair_plane = pairs_x * (uint32_t)g.tiles_y;
        std::vector<uint32_t> pv;
        for (uint32_t zc : {0u, nz / 2, nz - 1})
            for (uint32_t ty : {0u, (uint32_t)g.tiles_y - 1})
                for (uint32_t px : {0u, pairs_x - 1}) pv.push_back(zc * pair_plane + ty * pairs_x + px);
        pv.push_back(xcd_remap(pairs - 1, pairs));
        auto word = [&](size_t w) { pop(2 * w + 1, 6); };
        for (uint32_t pr : pv) {
            const int z = (int)(pr / pair_plane);
            const uint32_t r = pr - (uint32_t)z * pair_plane, ty = r / pairs_x, px = r - ty * pairs_x;
            const int X0 = (int)px * 2 * TILE_X;
            for (int wave = 0; wave < TILE_Y; ++wave) {
                const int y = std::min((int)ty * TILE_Y + wave, g.ny - 1);
                const uint32_t row_self = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
                bool lo_far = false, hi_far = false;
                int xh_lo = shift_axis(X0, 1, g.nx, bc[0], bc[1], lo_far);
                int xh_hi = shift_axis(X0 + 2 * TILE_X - 1, -1, g.nx, bc[0], bc[1], hi_far);
                if (lo_far) xh_lo = X0;
                if (hi_far) xh_hi = X0;
                for (int q = 0; q < 19; ++q) {
                    bool far = false;
                    const int ys = shift_axis(y, cy(q), g.ny, bc[2], bc[3], far);
                    const int zs = shift_axis(z, cz(q), g.nz, bc[4], bc[5], far);
                    const uint32_t row_src = far ? row_self : (uint32_t)ys + (uint32_t)g.ny * (uint32_t)zs;
                    const size_t base = pop_index<STORE_FP16_OFFSET>(g, q, 0, row_src);
                    if (base % 2 && !fail) fail = 6;
                    for (int lane : {0, 1, 31, 32, 63}) {
                        word((base + (size_t)X0) / 2 + (size_t)lane);
                        word(pop_index<STORE_FP16_OFFSET>(g, q, X0 + 2 * lane, row_self) >> 1);
                    }
                    if (cx(q) != 0) word((base + (size_t)(cx(q) > 0 ? xh_lo : xh_hi)) / 2);
                }
            }
        }
    }
    // the block -> tile remap
    std::vector<uint8_t> seen(g.tiles, 0);
    for (uint32_t b = 0; b < g.tiles; ++b) {
        const uint32_t t = xcd_remap(b, g.tiles);
        if (t >= g.tiles || seen[t]) {
            if (!fail) fail = 4;
            break;
        }
        seen[t] = 1;
    }
    out[0] = max_pop;
    out[1] = elements;
    out[2] = max_link;
    out[3] = table;
    out[4] = sources;
    return fail;
}

}  // extern "C"
