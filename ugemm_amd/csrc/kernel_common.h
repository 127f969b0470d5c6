// Device-side helpers shared by the GEMM kernel families.
#pragma once

#include <hip/hip_runtime.h>
#include <utility>

namespace ugemm {

// static_for<N>(f): f(std::integral_constant<int, i>{}) for i = 0 .. N-1,
// unrolled at compile time (indices usable in if constexpr and templates).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Maps a launch's linear block id to a (tile_m, tile_n) pair of C tiles.
// 1) XCD de-interleave: hardware hands consecutive block ids round-robin to
//    the 8 XCDs of gfx950; renumber so that each XCD (one L2) owns a
//    contiguous range of tile ids.
// 2) Grouped order: GROUP_M tile-rows are walked column by column so that the
//    blocks resident at one time share A row-panels and B column-panels.
// The mapping is a bijection of [0, tiles_m * tiles_n).
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int total = tiles_m * tiles_n;
    int pid = bid;
    if ((total & 7) == 0) pid = (bid & 7) * (total >> 3) + (bid >> 3);
    constexpr int GROUP_M = 8;
    const int per_group = GROUP_M * tiles_n;
    const int group = pid / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = pid - group * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

// Loads one operand tile (a C_EXT x S_EXT block: C_EXT elements along the
// operand's contiguous dimension, S_EXT along its strided dimension) from
// global memory into registers, VEC contiguous elements per chunk; VEC * sizeof(T)
// is 16 bytes (one 128-bit load) or VEC is 1 (scalar loads for unaligned data).
template <typename T, int VEC>
struct Vec16;
template <>
struct Vec16<float, 4> { typedef float4 type; };
template <>
struct Vec16<double, 2> { typedef double2 type; };

template <typename T, int C_EXT, int S_EXT, int NT, int VEC>
struct TileRegs {
    static constexpr int CHUNKS_C = C_EXT / VEC;
    static constexpr int NCHUNK = (C_EXT * S_EXT / VEC) / NT;
    static_assert((C_EXT * S_EXT / VEC) % NT == 0, "tile must split evenly over threads");
    static_assert(VEC == 1 || VEC * sizeof(T) == 16, "chunks are scalars or 16 bytes");
    T v[NCHUNK][VEC];

    __device__ __forceinline__ void load_vec(int i, const T *p) {
        if constexpr (VEC == 1) {
            v[i][0] = *p;
        } else {
            typedef typename Vec16<T, VEC>::type V;
            const V x = *reinterpret_cast<const V *>(p);
            const T *e = reinterpret_cast<const T *>(&x);
#pragma unroll
            for (int j = 0; j < VEC; ++j) v[i][j] = e[j];
        }
    }

    // base points at the tile origin; element (c, s) is base[s * ld + c].
    // cmax / smax: valid extents from the origin (only used when CHECK).
    // vec_ok == false (CHECK only): the operand is not 16-byte aligned, every
    // chunk is read with scalar loads.
    template <bool CHECK>
    __device__ __forceinline__ void load(const T *__restrict__ base, long ld, int t, int cmax, int smax,
                                         bool vec_ok = true) {
#pragma unroll
        for (int i = 0; i < NCHUNK; ++i) {
            const int f = t + i * NT;
            const int c = (f % CHUNKS_C) * VEC;
            const int s = f / CHUNKS_C;
            const T *p = base + (long)s * ld + c;
            if (!CHECK) {
                load_vec(i, p);
            } else {
                const bool srow = s < smax;
                if (vec_ok && srow && c + VEC <= cmax) {
                    load_vec(i, p);
                } else {
#pragma unroll
                    for (int j = 0; j < VEC; ++j) v[i][j] = (srow && c + j < cmax) ? p[j] : T(0);
                }
            }
        }
    }

    // Writes the registers into an LDS array L[k][x] of row length LD.
    // DIRECT: the contiguous dimension is x (the LDS fast index);
    // otherwise it is k and the chunk is written down a column.
    template <bool DIRECT, int LD>
    __device__ __forceinline__ void store(T *L, int t) const {
#pragma unroll
        for (int i = 0; i < NCHUNK; ++i) {
            const int f = t + i * NT;
            const int c = (f % CHUNKS_C) * VEC;
            const int s = f / CHUNKS_C;
            if constexpr (DIRECT) {
                if constexpr (VEC == 1) {
                    L[s * LD + c] = v[i][0];
                } else {
                    typedef typename Vec16<T, VEC>::type V;
                    V x;
                    T *e = reinterpret_cast<T *>(&x);
#pragma unroll
                    for (int j = 0; j < VEC; ++j) e[j] = v[i][j];
                    *reinterpret_cast<V *>(&L[s * LD + c]) = x;
                }
            } else {
#pragma unroll
                for (int j = 0; j < VEC; ++j) L[(c + j) * LD + s] = v[i][j];
            }
        }
    }
};

}  // namespace ugemm
