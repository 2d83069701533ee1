// shcfd_device.h -- device helpers shared by the kernel translation units.
#pragma once
#include "shcfd_internal.h"

namespace shcfd {

#define FLD(d, F, k, c) ((d).f[F][(long long)(k) * (d).NC + (c)])

template <int NY>
struct Dims {
    static constexpr int ny = NY;
    static constexpr int dq = NY + 3;
    static constexpr int dw = NY + 8;
    static constexpr int ig = dq + 1;  // gamma
    static constexpr int ih = dq + 2;  // H
    static constexpr int ir = dq + 3;  // R
    static constexpr int imu = dq + 4; // mu
};

// cell coordinates of this thread in a 64x4 tile
struct TileCoord {
    int b, i, j;
};
__device__ inline TileCoord tile_coord(const Tile* tiles) {
    Tile t = tiles[blockIdx.x];
    TileCoord c;
    c.b = t.b;
    c.i = t.i0 + (int)(threadIdx.x % TI);
    c.j = t.j0 + (int)(threadIdx.x / TI);
    return c;
}

__device__ inline double fsign1(double x) { return (x >= 0.0) ? 1.0 : -1.0; }
__device__ inline double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ inline double dmax(double a, double b) { return (a > b) ? a : b; }

// whole-wavefront shifts of a double by one lane (DPP wave_shr:1 / wave_shl:1);
// the lane that has no source receives 0.
__device__ inline double wave_shr1(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_shl1(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// agent-coherent 64-bit load/store (global_load/store ... sc1) used for
// data handed between workgroups inside one launch
__device__ inline void store_agent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_agent(const double* p) {
    unsigned long long v = __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}
__device__ inline void store_flag(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int load_flag(const int* p) {
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every vector-memory operation of this wave has completed
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// quad broadcast: every lane of a 4-lane group gets lane k of the group
template <int K>
__device__ inline double quad_bcast(double x) {
    constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
    int lo = __double2loint(x), hi = __double2hiint(x);
    // every lane has a source within its quad: bound_ctrl set, no old value
    lo = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// lane L receives lane L-4 (0 for L < 4) / lane L+4 (0 for L > 59)
__device__ inline double shift_up4(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, true);
        hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, true);
    }
    return __hiloint2double(hi, lo);
}
__device__ inline double shift_down4(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, true);
        hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, true);
    }
    return __hiloint2double(hi, lo);
}

// quad broadcast with the source lane given at run time (constant after
// unrolling, so the switch folds to one DPP pair)
__device__ inline double quad_bcast_n(double x, int k) {
    switch (k) {
    case 0: return quad_bcast<0>(x);
    case 1: return quad_bcast<1>(x);
    case 2: return quad_bcast<2>(x);
    default: return quad_bcast<3>(x);
    }
}

}  // namespace shcfd
