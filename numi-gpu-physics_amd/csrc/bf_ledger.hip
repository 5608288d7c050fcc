// Momentum ledger of a completed step, in FP64 on the device.
//
// Reference: Sources/BirdFlowMetal/Metal/BirdFlow.metal
//   measureFluidMassMomentum :4272-4317, measureFluidMassMomentumChange
//   :4319-4357, captureExternalFluidSources :4359-4526; driven by
//   BirdFlowSimulation.advanceWithCoupledMomentumLedger
//   (Sources/BirdFlowMetal/BirdFlowSimulation.swift:805-1085).
//
// The reference sums these in FP32; here every per-cell term is formed in
// FP64 from the stored populations and summed in a fixed order (wavefront
// butterflies, four waves per block, blocks in index order), so a ledger is
// reproducible run to run.  Sums (mass, momentum x y z):
//   [0]  change:  fluid populations after the step (current mask) minus
//                 before it (previous mask), masking each side on its own;
//   [4]  open boundaries: per link pulled from beyond a far-field face,
//        (f_far - f_out, (f_far + f_out) c) with f_out = f_in,opp(q) of the
//        cell; a zero-gradient (outlet) face contributes the same with the
//        clamped source's population as the incoming one;
//   [8]  sponge: s (f_far - f_actual) / (1 - s) per direction of each
//        sponge cell (f_actual = (1 - s) f_post + s f_far);
//   [12] persistent wall links (target fluid in both steps, source solid
//        in both): halfway moving bounce-back exchange with the wall
//        density far_rho, 2 w_q rho (c.u_wall) / cs^2.
// Counts: open-boundary links, sponge cells, persistent links, cells whose
// occupancy changed.
#include <mutex>

#include "bf_fluid.h"

namespace bf {

namespace {

constexpr int LEDGER_SUMS = 16;
constexpr int LEDGER_COUNTS = 4;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block partial of NV sums and NC counts: wave butterflies, then the four
// waves in order.
template <int NV, int NC>
__device__ __forceinline__ void block_partial(const double* v, const uint32_t* n, double* out_v,
                                              unsigned long long* out_n) {
    __shared__ double sv[4][NV];
    __shared__ uint32_t sn[4][NC > 0 ? NC : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) sv[wave][i] = s;
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const uint32_t s = wave_sum_u(n[i]);
        if (lane == 0) sn[wave][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int i = threadIdx.x;
        out_v[(size_t)blockIdx.x * NV + i] = ((sv[0][i] + sv[1][i]) + sv[2][i]) + sv[3][i];
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + NC) {
        const int i = threadIdx.x - 64;
        out_n[(size_t)blockIdx.x * NC + i] = (unsigned long long)sn[0][i] + sn[1][i] + sn[2][i] + sn[3][i];
    }
}

__device__ __forceinline__ void add4(double* a, double m, double dir_x, double dir_y, double dir_z, double mom) {
    a[0] += m;
    a[1] += mom * dir_x;
    a[2] += mom * dir_y;
    a[3] += mom * dir_z;
}

}  // namespace

// Mass and momentum of the fluid cells (mask 0) of one lattice.
template <int S>
__global__ __launch_bounds__(256) void ledger_fluid(const typename Store<S>::T* __restrict__ f,
                                                    const uint8_t* __restrict__ solid, Grid g,
                                                    double* __restrict__ out_v) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t gid = blockIdx.x * 256u + threadIdx.x; gid < g.n; gid += gridDim.x * 256u) {
        if (solid[gid] != 0) continue;
        const uint32_t row = gid / (uint32_t)g.nx;
        const int x = (int)(gid - row * (uint32_t)g.nx);
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const double fq = (double)Store<S>::load(f, pop_index<S>(g, q, x, row), q);
            add4(v, fq, (double)cx(q), (double)cy(q), (double)cz(q), fq);
        }
    }
    block_partial<4, 0>(v, nullptr, out_v, nullptr);
}

template <int S>
__global__ __launch_bounds__(256) void ledger_sources(const typename Store<S>::T* __restrict__ fin,
                                                      const typename Store<S>::T* __restrict__ fout,
                                                      const uint8_t* __restrict__ solid_prev,
                                                      const uint8_t* __restrict__ solid_cur,
                                                      const float4* __restrict__ wall, StepParams p,
                                                      double* __restrict__ out_v,
                                                      unsigned long long* __restrict__ out_n) {
    const Grid& g = p.g;
    double v[LEDGER_SUMS];
#pragma unroll
    for (int i = 0; i < LEDGER_SUMS; ++i) v[i] = 0.0;
    uint32_t n[LEDGER_COUNTS] = {0u, 0u, 0u, 0u};
    for (uint32_t gid = blockIdx.x * 256u + threadIdx.x; gid < g.n; gid += gridDim.x * 256u) {
        const uint32_t row = gid / (uint32_t)g.nx;
        const int x = (int)(gid - row * (uint32_t)g.nx);
        const int y = (int)(row % (uint32_t)g.ny), z = (int)(row / (uint32_t)g.ny);
        const bool was = solid_prev[gid] != 0, is = solid_cur[gid] != 0;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const size_t i = pop_index<S>(g, q, x, row);
            double d = 0.0;
            if (!is) d += (double)Store<S>::load(fout, i, q);
            if (!was) d -= (double)Store<S>::load(fin, i, q);
            add4(v, d, (double)cx(q), (double)cy(q), (double)cz(q), d);
        }
        if (was != is) n[3] += 1u;
        if (is) continue;
        for (int q = 1; q < 19; ++q) {
            bool far = false;
            const int ux = x - cx(q), uy = y - cy(q), uz = z - cz(q);
            const int sx = shift_axis(x, cx(q), g.nx, p.bc[0][0], p.bc[0][1], far);
            const int sy = shift_axis(y, cy(q), g.ny, p.bc[1][0], p.bc[1][1], far);
            const int sz = shift_axis(z, cz(q), g.nz, p.bc[2][0], p.bc[2][1], far);
            const bool open = (ux < 0 && p.bc[0][0] == BC_OUTLET) || (ux >= g.nx && p.bc[0][1] == BC_OUTLET) ||
                              (uy < 0 && p.bc[1][0] == BC_OUTLET) || (uy >= g.ny && p.bc[1][1] == BC_OUTLET) ||
                              (uz < 0 && p.bc[2][0] == BC_OUTLET) || (uz >= g.nz && p.bc[2][1] == BC_OUTLET);
            const double dxq = (double)cx(q), dyq = (double)cy(q), dzq = (double)cz(q);
            const double outgoing = (double)Store<S>::load(fin, pop_index<S>(g, opp(q), x, row), opp(q));
            if (far || open) {
                const uint32_t srow = (uint32_t)sy + (uint32_t)g.ny * (uint32_t)sz;
                const double incoming =
                    far ? (double)p.far_eq[q] : (double)Store<S>::load(fin, pop_index<S>(g, q, sx, srow), q);
                add4(v + 4, incoming - outgoing, dxq, dyq, dzq, incoming + outgoing);
                n[0] += 1u;
                continue;
            }
            const uint32_t src = (uint32_t)sx + (uint32_t)g.nx * ((uint32_t)sy + (uint32_t)g.ny * (uint32_t)sz);
            if (!was && solid_cur[src] != 0 && solid_prev[src] != 0) {
                const float4 u = wall[src];
                const double cu = dxq * (double)u.x + dyq * (double)u.y + dzq * (double)u.z;
                const double corr = 2.0 * (double)weight<float>(q) * (double)p.far_rho * cu * 3.0;
                add4(v + 12, corr, dxq, dyq, dzq, 2.0 * outgoing + corr);
                n[2] += 1u;
            }
        }
        if (p.sponge_strength > 0.0f) {
            const float s = sponge_factor(p, x, y, z);
            if (s > 0.0f) {
                const double sd = (double)s;
                const double inv = 1.0 / fmax(1.0 - sd, 1.0e-7);
#pragma unroll
                for (int q = 0; q < 19; ++q) {
                    const double actual = (double)Store<S>::load(fout, pop_index<S>(g, q, x, row), q);
                    const double d = sd * ((double)p.far_eq[q] - actual) * inv;
                    add4(v + 8, d, (double)cx(q), (double)cy(q), (double)cz(q), d);
                }
                n[1] += 1u;
            }
        }
    }
    block_partial<LEDGER_SUMS, LEDGER_COUNTS>(v, n, out_v, out_n);
}

// Fixed-order sum of the block partials: lane i adds value i of every block
// in block order.
__global__ void ledger_finish(const double* __restrict__ pv, const unsigned long long* __restrict__ pn,
                              uint32_t blocks, int nv, int nc, double* __restrict__ out_v,
                              unsigned long long* __restrict__ out_n) {
    const int i = (int)threadIdx.x;
    if (i < nv) {
        double s = 0.0;
        for (uint32_t b = 0; b < blocks; ++b) s += pv[(size_t)b * nv + i];
        out_v[i] = s;
    } else if (i >= 64 && i < 64 + nc) {
        unsigned long long s = 0;
        for (uint32_t b = 0; b < blocks; ++b) s += pn[(size_t)b * nc + (i - 64)];
        out_n[i - 64] = s;
    }
}

template <int S>
static hipError_t ledger_s(const LedgerBuffers& b, const StepParams& p, bool sources, hipStream_t s) {
    using T = typename Store<S>::T;
    if (sources) {
        hipLaunchKernelGGL(ledger_sources<S>, dim3(b.blocks), dim3(256), 0, s, (const T*)b.f_in, (const T*)b.f_out,
                           b.solid_prev, b.solid_cur, b.wall, p, b.partial_v, b.partial_n);
        hipLaunchKernelGGL(ledger_finish, dim3(1), dim3(128), 0, s, (const double*)b.partial_v,
                           (const unsigned long long*)b.partial_n, b.blocks, LEDGER_SUMS, LEDGER_COUNTS, b.out_v,
                           b.out_n);
    } else {
        hipLaunchKernelGGL(ledger_fluid<S>, dim3(b.blocks), dim3(256), 0, s, (const T*)b.f_out, b.solid_cur, p.g,
                           b.partial_v);
        hipLaunchKernelGGL(ledger_finish, dim3(1), dim3(128), 0, s, (const double*)b.partial_v,
                           (const unsigned long long*)nullptr, b.blocks, 4, 0, b.out_v, b.out_n);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation ledger entry point
// (bf_ledger_step_sources_noalloc): the block partials and the reduced sums
// live in a file-scope device array instead of buffers the context allocates
// on its first ledger call.  Every GPU holds its own copy of the array until
// the process ends, at its own address, looked up once per device
// (hipGetSymbolAddress, with that device current) and kept per device; host
// code uses only that address.  All contexts on one GPU share the array, so
// calls that overlap on one GPU race.  Nothing in it needs clearing: each
// launch writes every partial it later reads, and the finishing kernel
// writes the outputs.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_V = (size_t)LEDGER_SCRATCH_BLOCKS * LEDGER_SUMS * sizeof(double);
constexpr size_t SCRATCH_N = (size_t)LEDGER_SCRATCH_BLOCKS * LEDGER_COUNTS * sizeof(unsigned long long);
constexpr size_t SCRATCH_OUT_V = SCRATCH_V + SCRATCH_N;
constexpr size_t SCRATCH_OUT_N = SCRATCH_OUT_V + LEDGER_SUMS * sizeof(double);
constexpr size_t SCRATCH_BYTES = (SCRATCH_OUT_N + LEDGER_COUNTS * sizeof(unsigned long long) + 255) / 256 * 256;
static_assert(SCRATCH_BYTES <= 1u << 20, "the ledger scratch stays under 1 MiB per GPU");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mutex;
unsigned char* scratch_address[SCRATCH_DEVICES] = {};
}  // namespace

hipError_t ledger_scratch_setup(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_address[device]) return hipSuccess;
    unsigned char* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    scratch_address[device] = ptr;
    return hipSuccess;
}

hipError_t ledger_scratch_buffers(int device, uint32_t blocks, LedgerBuffers& b) {
    if (blocks == 0 || blocks > LEDGER_SCRATCH_BLOCKS) return hipErrorInvalidValue;
    if (device < 0 || device >= SCRATCH_DEVICES) return hipErrorInvalidDevice;
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(scratch_mutex);
        base = scratch_address[device];
    }
    if (!base) return hipErrorNotInitialized;
    b.blocks = blocks;
    b.partial_v = (double*)base;
    b.partial_n = (unsigned long long*)(base + SCRATCH_V);
    b.out_v = (double*)(base + SCRATCH_OUT_V);
    b.out_n = (unsigned long long*)(base + SCRATCH_OUT_N);
    return hipSuccess;
}

hipError_t launch_ledger(int storage, const LedgerBuffers& b, const StepParams& p, bool sources, hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return ledger_s<STORE_FP32>(b, p, sources, s);
        case STORE_FP16_OFFSET: return ledger_s<STORE_FP16_OFFSET>(b, p, sources, s);
        case STORE_BF16_OFFSET: return ledger_s<STORE_BF16_OFFSET>(b, p, sources, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bf
