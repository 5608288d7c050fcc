// Non-templated fluid kernels and their launchers: the general-tile work
// list, the fixed-order load reduction, initialisation, field extraction and
// storage conversion; plus the dispatch of the fluid step over arithmetic,
// storage and collision (instantiations live in bf_fluid_fused.hip and
// bf_fluid_strict.hip).
#include "bf_fluid.h"

namespace bf {

// Work list of the tiles that take the general path this step: flagged by
// the static mask (partial or sponge tiles) or by solid cells of the
// previous / current geometry.  Appends are aggregated per wavefront.
__global__ __launch_bounds__(256) void build_general_list(const uint8_t* __restrict__ static_flags,
                                                          const uint8_t* __restrict__ dyn_prev,
                                                          const uint8_t* __restrict__ dyn_cur, uint32_t tiles,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool flagged = t < tiles && (static_flags[t] | dyn_prev[t] | dyn_cur[t]) != 0;
    const unsigned long long b = __ballot(flagged);
    if (b == 0) return;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(b));
    base = __shfl(base, 0, 64);
    if (flagged)
        list[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = t;
}



// Fixed-order sum of the per-tile partials of the tiles the general kernel
// processed (flags identify them; other tiles contribute exactly zero).
__global__ __launch_bounds__(256) void reduce_loads_tiles(const double* __restrict__ partials,
                                                          const uint8_t* __restrict__ static_flags,
                                                          const uint8_t* __restrict__ dyn_prev,
                                                          const uint8_t* __restrict__ dyn_cur, uint32_t tiles,
                                                          double* __restrict__ block_out) {
    constexpr uint32_t PER_THREAD = 4;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * PER_THREAD;
    for (uint32_t k = 0; k < PER_THREAD; ++k) {
        const uint32_t t = base + k;
        if (t < tiles && (static_flags[t] | dyn_prev[t] | dyn_cur[t]) != 0) {
#pragma unroll
            for (int i = 0; i < 7; ++i) v[i] += partials[(size_t)t * 8 + i];
            v[7] = fmax(v[7], partials[(size_t)t * 8 + 7]);
        }
    }
    block_reduce8(v, block_out + (size_t)blockIdx.x * 8);
}

// Adds the fast tiles' limiter activations (an integer count and a maximum,
// exact in any order) and clears them for the step that reuses the slot.
__global__ __launch_bounds__(256) void reduce_loads_final(const double* __restrict__ block_out, uint32_t blocks,
                                                          uint32_t* __restrict__ activation,
                                                          double* __restrict__ last, double* __restrict__ history,
                                                          uint32_t history_slot) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
#pragma unroll
        for (int i = 0; i < 7; ++i) v[i] += block_out[(size_t)b * 8 + i];
        v[7] = fmax(v[7], block_out[(size_t)b * 8 + 7]);
    }
    __shared__ double total[8];
    block_reduce8(v, total);
    if (threadIdx.x < 8) {
        double t = total[threadIdx.x];
        if (threadIdx.x == 6) t += (double)activation[0];
        if (threadIdx.x == 7) t = fmax(t, (double)__uint_as_float(activation[1]));
        last[threadIdx.x] = t;
        if (history) history[(size_t)history_slot * 8 + threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < 2) activation[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Initialisation, readback and storage conversion.
// ---------------------------------------------------------------------------
template <int S>
__global__ void init_equilibrium(typename Store<S>::T* __restrict__ f, const uint8_t* __restrict__ solid,
                                 const float4* __restrict__ wall, float fux, float fuy, float fuz, Grid g) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n) return;
    const uint32_t row = gid / (uint32_t)g.nx;
    const int x = (int)(gid - row * (uint32_t)g.nx);
    float ux = fux, uy = fuy, uz = fuz;
    if (solid[gid] != 0) {
        const float4 w = wall[gid];
        ux = w.x;
        uy = w.y;
        uz = w.z;
    }
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        Store<S>::store(f, pop_index<S>(g, q, x, row), q, equilibrium<float>(q, 1.0f, ux, uy, uz));
    });
}

// Shear wave u_x = A sin(2 pi y / ny) (BirdFlow.metal:2132-2162); the sine
// of the FP32 phase is evaluated in FP64 and rounded once.
template <int S>
__global__ void init_shear_wave(typename Store<S>::T* __restrict__ f, float amplitude, Grid g) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n) return;
    const uint32_t row = gid / (uint32_t)g.nx;
    const int x = (int)(gid - row * (uint32_t)g.nx);
    const uint32_t y = row % (uint32_t)g.ny;
    const float phase = 2.0f * 3.14159265358979323846f * (float)y / (float)g.ny;
    const float ux = amplitude * (float)sin((double)phase);
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        Store<S>::store(f, pop_index<S>(g, q, x, row), q, equilibrium<float>(q, 1.0f, ux, 0.0f, 0.0f));
    });
}

template <int S>
__global__ void extract_fields(const typename Store<S>::T* __restrict__ f, float* __restrict__ density,
                               float* __restrict__ velocity3, Grid g) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n) return;
    const uint32_t row = gid / (uint32_t)g.nx;
    const int x = (int)(gid - row * (uint32_t)g.nx);
    float v[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        v[q] = Store<S>::load(f, pop_index<S>(g, q, x, row), q);
    });
    float rho, mx, my, mz;
    moments(v, rho, mx, my, mz);
    rho = fmaxf(rho, 1.0e-8f);
    density[gid] = rho;
    velocity3[3 * (size_t)gid + 0] = mx / rho;
    velocity3[3 * (size_t)gid + 1] = my / rho;
    velocity3[3 * (size_t)gid + 2] = mz / rho;
}

// Host layout (direction-major, float) <-> device layout and storage.
template <int S>
__global__ void pack_populations(const float* __restrict__ src, typename Store<S>::T* __restrict__ dst, size_t count,
                                 Grid g) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int q = (int)(i / g.n);
    const uint32_t gid = (uint32_t)(i - (size_t)q * g.n);
    const uint32_t row = gid / (uint32_t)g.nx;
    Store<S>::store(dst, pop_index<S>(g, q, (int)(gid - row * (uint32_t)g.nx), row), q, src[i]);
}

template <int S>
__global__ void unpack_populations(const typename Store<S>::T* __restrict__ src, float* __restrict__ dst,
                                   size_t count, Grid g) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int q = (int)(i / g.n);
    const uint32_t gid = (uint32_t)(i - (size_t)q * g.n);
    const uint32_t row = gid / (uint32_t)g.nx;
    dst[i] = Store<S>::load(src, pop_index<S>(g, q, (int)(gid - row * (uint32_t)g.nx), row), q);
}

// Stored values moved unchanged between the host's direction-major order
// ([q][z][y][x]) and the device's row-interleaved one (checkpoint images of
// the storage form, bit-exact for every storage).
template <typename T>
__global__ void permute_populations(const T* __restrict__ src, T* __restrict__ dst, size_t count, Grid g,
                                    int to_device) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int q = (int)(i / g.n);
    const uint32_t gid = (uint32_t)(i - (size_t)q * g.n);
    const uint32_t row = gid / (uint32_t)g.nx;
    const size_t d = pop_index<STORE_FP32>(g, q, (int)(gid - row * (uint32_t)g.nx), row);
    if (to_device)
        dst[d] = src[i];
    else
        dst[i] = src[d];
}

// reducePopulationMinimum (BirdFlow.metal:6404-6459) over the stored lattice
// in the host's direction-major numbering (i = q N + cell): the smallest
// value, a non-finite value comparing as -infinity, ties to the smallest
// index.  That order is total, so the result does not depend on how the
// elements are split among lanes and blocks: each lane keeps the best of its
// grid-stride elements, a workgroup tree keeps the best of the lanes, one
// block the best of the workgroups.
__device__ __forceinline__ bool pop_better(float ca, unsigned long long ia, float cb, unsigned long long ib) {
    return ca < cb || (ca == cb && ia < ib);
}

__device__ void pop_min_block(PopulationMinimum best, PopulationMinimum* __restrict__ out) {
    __shared__ float cmp[256], raw[256];
    __shared__ unsigned long long idx[256];
    __shared__ uint32_t bad[256];
    const uint32_t t = threadIdx.x;
    cmp[t] = best.comparison;
    raw[t] = best.value;
    idx[t] = best.index;
    bad[t] = best.non_finite;
    __syncthreads();
    for (uint32_t stride = 128; stride > 0; stride >>= 1) {
        if (t < stride && pop_better(cmp[t + stride], idx[t + stride], cmp[t], idx[t])) {
            cmp[t] = cmp[t + stride];
            raw[t] = raw[t + stride];
            idx[t] = idx[t + stride];
            bad[t] = bad[t + stride];
        }
        __syncthreads();
    }
    if (t == 0) *out = PopulationMinimum{cmp[0], raw[0], idx[0], bad[0], 0u};
}

template <int S>
__global__ __launch_bounds__(256) void population_minimum(const typename Store<S>::T* __restrict__ f, Grid g,
                                                          PopulationMinimum* __restrict__ partials) {
    const size_t total = 19ull * g.n;
    PopulationMinimum best{INFINITY, INFINITY, ~0ull, 0u, 0u};
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        // device position e = (row 19 + q) nx + x
        const size_t r = e / (size_t)g.nx;
        const int x = (int)(e - r * (size_t)g.nx);
        const int q = (int)(r % 19u);
        const size_t row = r / 19u;
        const float v = Store<S>::load(f, e, q);
        const bool finite = isfinite(v);
        const float c = finite ? v : -INFINITY;
        const unsigned long long i = (unsigned long long)q * g.n + (unsigned long long)x + (unsigned long long)g.nx * row;
        if (pop_better(c, i, best.comparison, best.index)) best = PopulationMinimum{c, v, i, finite ? 0u : 1u, 0u};
    }
    pop_min_block(best, partials + blockIdx.x);
}

__global__ __launch_bounds__(256) void population_minimum_final(const PopulationMinimum* __restrict__ partials,
                                                                uint32_t count, PopulationMinimum* __restrict__ out) {
    PopulationMinimum best{INFINITY, INFINITY, ~0ull, 0u, 0u};
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x)
        if (pop_better(partials[i].comparison, partials[i].index, best.comparison, best.index)) best = partials[i];
    pop_min_block(best, out);
}

// One stored population set to `value` (the host's direction-major index):
// fault injection for the safety checks.
template <int S>
__global__ void poke_population(typename Store<S>::T* __restrict__ f, Grid g, unsigned long long index, float value) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int q = (int)(index / g.n);
    const uint32_t gid = (uint32_t)(index - (unsigned long long)q * g.n);
    const uint32_t row = gid / (uint32_t)g.nx;
    Store<S>::store(f, pop_index<S>(g, q, (int)(gid - row * (uint32_t)g.nx), row), q, value);
}

template <int S>
static hipError_t pop_min_s(const void* f, const Grid& g, PopulationMinimum* partials, uint32_t blocks,
                            PopulationMinimum* out, hipStream_t s) {
    hipLaunchKernelGGL(population_minimum<S>, dim3(blocks), dim3(256), 0, s, (const typename Store<S>::T*)f, g,
                       partials);
    hipLaunchKernelGGL(population_minimum_final, dim3(1), dim3(256), 0, s, (const PopulationMinimum*)partials, blocks,
                       out);
    return hipGetLastError();
}

hipError_t launch_population_minimum(int storage, const void* f, const Grid& g, PopulationMinimum* partials,
                                     uint32_t blocks, PopulationMinimum* out, hipStream_t s) {
    if (blocks < 1 || blocks > POPULATION_MINIMUM_BLOCKS) return hipErrorInvalidValue;
    switch (storage) {
        case STORE_FP32: return pop_min_s<STORE_FP32>(f, g, partials, blocks, out, s);
        case STORE_FP16_OFFSET: return pop_min_s<STORE_FP16_OFFSET>(f, g, partials, blocks, out, s);
        case STORE_BF16_OFFSET: return pop_min_s<STORE_BF16_OFFSET>(f, g, partials, blocks, out, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_poke_population(int storage, void* f, const Grid& g, unsigned long long index, float value,
                                  hipStream_t s) {
    if (index >= 19ull * g.n) return hipErrorInvalidValue;
    switch (storage) {
        case STORE_FP32:
            hipLaunchKernelGGL(poke_population<STORE_FP32>, dim3(1), dim3(64), 0, s, (float*)f, g, index, value);
            break;
        case STORE_FP16_OFFSET:
            hipLaunchKernelGGL(poke_population<STORE_FP16_OFFSET>, dim3(1), dim3(64), 0, s, (_Float16*)f, g, index,
                               value);
            break;
        case STORE_BF16_OFFSET:
            hipLaunchKernelGGL(poke_population<STORE_BF16_OFFSET>, dim3(1), dim3(64), 0, s, (uint16_t*)f, g, index,
                               value);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_load_reduction(const FluidBuffers& b, const StepParams& p, uint32_t history_slot, hipStream_t s) {
    if (!p.accumulate_loads) return hipSuccess;
    const uint32_t blocks = (p.g.tiles + 1023) / 1024;
    hipLaunchKernelGGL(reduce_loads_tiles, dim3(blocks), dim3(256), 0, s, b.partials, b.static_flags, b.dyn_prev,
                       b.dyn_cur, p.g.tiles, b.block_out);
    hipLaunchKernelGGL(reduce_loads_final, dim3(1), dim3(256), 0, s, b.block_out, blocks, b.activation, b.last_load,
                       b.history, history_slot);
    return hipGetLastError();
}

hipError_t launch_fluid_fast(int arith, int storage, int coll, const FluidBuffers& b, const StepParams& p,
                             hipStream_t s) {
    if (arith == ARITH_STRICT) {
        if (storage != STORE_FP32) return hipErrorInvalidValue;
        return launch_fluid_fast_strict(coll, b, p, s);
    }
    return launch_fluid_fast_fused(storage, coll, b, p, s);
}

hipError_t launch_fluid_general(int arith, int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                int general_blocks, hipStream_t s) {
    if (arith == ARITH_STRICT) {
        if (storage != STORE_FP32) return hipErrorInvalidValue;
        return launch_fluid_general_strict(coll, b, p, general_blocks, s);
    }
    return launch_fluid_general_fused(storage, coll, b, p, general_blocks, s);
}

hipError_t launch_build_general_list(const uint8_t* static_flags, const uint8_t* dyn_prev, const uint8_t* dyn_cur,
                                     uint32_t tiles, uint32_t* list, uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(build_general_list, dim3((tiles + 255) / 256), dim3(256), 0, s, static_flags, dyn_prev,
                       dyn_cur, tiles, list, count);
    return hipGetLastError();
}

template <int S>
static hipError_t init_eq_s(void* f, const uint8_t* solid, const float4* wall, const float* fu, const Grid& g,
                            hipStream_t s) {
    hipLaunchKernelGGL(init_equilibrium<S>, dim3((g.n + 255) / 256), dim3(256), 0, s, (typename Store<S>::T*)f, solid,
                       wall, fu[0], fu[1], fu[2], g);
    return hipGetLastError();
}

hipError_t launch_init_equilibrium(int storage, void* f, const uint8_t* solid, const float4* wall,
                                   const float* far_u, const Grid& g, hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return init_eq_s<STORE_FP32>(f, solid, wall, far_u, g, s);
        case STORE_FP16_OFFSET: return init_eq_s<STORE_FP16_OFFSET>(f, solid, wall, far_u, g, s);
        case STORE_BF16_OFFSET: return init_eq_s<STORE_BF16_OFFSET>(f, solid, wall, far_u, g, s);
        default: return hipErrorInvalidValue;
    }
}

template <int S>
static hipError_t init_sw_s(void* f, float a, const Grid& g, hipStream_t s) {
    hipLaunchKernelGGL(init_shear_wave<S>, dim3((g.n + 255) / 256), dim3(256), 0, s, (typename Store<S>::T*)f, a, g);
    return hipGetLastError();
}

hipError_t launch_init_shear_wave(int storage, void* f, float amplitude, const Grid& g, hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return init_sw_s<STORE_FP32>(f, amplitude, g, s);
        case STORE_FP16_OFFSET: return init_sw_s<STORE_FP16_OFFSET>(f, amplitude, g, s);
        case STORE_BF16_OFFSET: return init_sw_s<STORE_BF16_OFFSET>(f, amplitude, g, s);
        default: return hipErrorInvalidValue;
    }
}

template <int S>
static hipError_t fields_s(const void* f, float* rho, float* u, const Grid& g, hipStream_t s) {
    hipLaunchKernelGGL(extract_fields<S>, dim3((g.n + 255) / 256), dim3(256), 0, s, (const typename Store<S>::T*)f,
                       rho, u, g);
    return hipGetLastError();
}

hipError_t launch_extract_fields(int storage, const void* f, float* rho, float* u, const Grid& g, hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return fields_s<STORE_FP32>(f, rho, u, g, s);
        case STORE_FP16_OFFSET: return fields_s<STORE_FP16_OFFSET>(f, rho, u, g, s);
        case STORE_BF16_OFFSET: return fields_s<STORE_BF16_OFFSET>(f, rho, u, g, s);
        default: return hipErrorInvalidValue;
    }
}

template <int S>
static hipError_t conv_s(bool pack, const void* src, void* dst, size_t count, const Grid& g, hipStream_t s) {
    const unsigned blocks = (unsigned)((count + 255) / 256);
    if (pack)
        hipLaunchKernelGGL(pack_populations<S>, dim3(blocks), dim3(256), 0, s, (const float*)src,
                           (typename Store<S>::T*)dst, count, g);
    else
        hipLaunchKernelGGL(unpack_populations<S>, dim3(blocks), dim3(256), 0, s, (const typename Store<S>::T*)src,
                           (float*)dst, count, g);
    return hipGetLastError();
}

hipError_t launch_permute_populations(int storage, bool to_device, const void* src, void* dst, size_t count,
                                      const Grid& g, hipStream_t s) {
    const unsigned blocks = (unsigned)((count + 255) / 256);
    if (storage == STORE_FP32)
        hipLaunchKernelGGL(permute_populations<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t*)src,
                           (uint32_t*)dst, count, g, to_device ? 1 : 0);
    else if (storage == STORE_FP16_OFFSET || storage == STORE_BF16_OFFSET)
        hipLaunchKernelGGL(permute_populations<uint16_t>, dim3(blocks), dim3(256), 0, s, (const uint16_t*)src,
                           (uint16_t*)dst, count, g, to_device ? 1 : 0);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_convert_populations(int storage, bool pack, const void* src, void* dst, size_t count, const Grid& g,
                                      hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return conv_s<STORE_FP32>(pack, src, dst, count, g, s);
        case STORE_FP16_OFFSET: return conv_s<STORE_FP16_OFFSET>(pack, src, dst, count, g, s);
        case STORE_BF16_OFFSET: return conv_s<STORE_BF16_OFFSET>(pack, src, dst, count, g, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bf
