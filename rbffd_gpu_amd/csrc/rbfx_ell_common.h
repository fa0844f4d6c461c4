// Device layout of a fused ELL operator set and the gather-dot core shared by the SpMV,
// SpMM and RK4 stage kernels.
//
// Layout in HBM (n_pad = n_rows rounded up to ELL_ROW_PAD):
//   idx_t[j * n_pad + row]                 int32, slot-major: one coalesced load per slot
//   w_t[(j * n_pad + row) * n_ops + o]     weights of the n_ops operators of one (slot,row)
//                                          side by side, so one vector load fetches all of
//                                          them; FP64 or FP32 storage
// Fields are node-major u[node * F + f], so the gathered neighbour record of FPT fields is
// one contiguous vector load. Pad rows carry index 0 and weight 0.
#pragma once
#include "rbfx_internal.h"

#define ELL_ROW_PAD 256

struct rbfx_ell {
    rbfx_ctx* ctx = nullptr;
    int32_t n_rows = 0, n_cols = 0, width = 0, n_ops = 0, storage = RBFX_F64;
    int32_t n_pad = 0;
    uint32_t types[16] = {0};
    int32_t* idx_t = nullptr;  // device
    void* w_t = nullptr;       // device, double or float
};

template <typename T, int N>
struct alignas(((N * sizeof(T)) & (N * sizeof(T) - 1)) == 0 && N * sizeof(T) <= 16 ? N * sizeof(T) : sizeof(T)) Pack {
    T v[N];
};

// Loads N consecutive T from p (alignment of the Pack type is guaranteed by the caller).
template <typename T, int N>
__device__ __forceinline__ void load_pack(const T* __restrict__ p, T (&out)[N]) {
    constexpr int B = (int)(N * sizeof(T));
    if constexpr (B <= 16 && (B & (B - 1)) == 0) {
        Pack<T, N> q = *reinterpret_cast<const Pack<T, N>*>(p);
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = q.v[i];
    } else if constexpr (B % 16 == 0) {
        constexpr int C = 16 / sizeof(T);
#pragma unroll
        for (int c = 0; c < N / C; c++) {
            Pack<T, C> q = *reinterpret_cast<const Pack<T, C>*>(p + c * C);
#pragma unroll
            for (int i = 0; i < C; i++) out[c * C + i] = q.v[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = p[i];
    }
}

template <typename T, int N>
__device__ __forceinline__ void store_pack(T* __restrict__ p, const T (&in)[N]) {
    constexpr int B = (int)(N * sizeof(T));
    if constexpr (B <= 16 && (B & (B - 1)) == 0) {
        Pack<T, N> q;
#pragma unroll
        for (int i = 0; i < N; i++) q.v[i] = in[i];
        *reinterpret_cast<Pack<T, N>*>(p) = q;
    } else if constexpr (B % 16 == 0) {
        constexpr int C = 16 / sizeof(T);
#pragma unroll
        for (int c = 0; c < N / C; c++) {
            Pack<T, C> q;
#pragma unroll
            for (int i = 0; i < C; i++) q.v[i] = in[c * C + i];
            *reinterpret_cast<Pack<T, C>*>(p + c * C) = q;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) p[i] = in[i];
    }
}

// Weights of all NOPS operators at one (slot,row): NOPS == 3 is loaded as 3 scalars.
template <typename WT, int NOPS>
__device__ __forceinline__ void load_w(const WT* __restrict__ p, WT (&w)[NOPS]) {
    if constexpr (NOPS == 3) {
        w[0] = p[0]; w[1] = p[1]; w[2] = p[2];
    } else {
        load_pack<WT, NOPS>(p, w);
    }
}

template <typename AT>
__device__ __forceinline__ AT fma_acc(AT a, AT b, AT c);
template <>
__device__ __forceinline__ double fma_acc<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_acc<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// acc[o][f] = sum_j w_o[row][j] * u[idx[row][j]*F + fofs + f], j ascending.
template <typename WT, typename UT, typename AT, int NOPS, int FPT>
__device__ __forceinline__ void ell_gather_dot(const int32_t* __restrict__ idx_t, const WT* __restrict__ w_t,
                                               const UT* __restrict__ u, int row, int n_pad, int width, int F,
                                               int fofs, AT (&acc)[NOPS][FPT]) {
#pragma unroll
    for (int o = 0; o < NOPS; o++)
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[o][f] = (AT)0;
    const int32_t* ip = idx_t + row;
    const WT* wp = w_t + (size_t)row * NOPS;
    const size_t wstride = (size_t)n_pad * NOPS;
    int j = 0;
    // two slots per iteration keeps two independent gathers in flight per thread
    for (; j + 1 < width; j += 2) {
        int n0 = __builtin_nontemporal_load(ip + (size_t)j * n_pad);
        int n1 = __builtin_nontemporal_load(ip + (size_t)(j + 1) * n_pad);
        WT w0[NOPS], w1[NOPS];
        load_w<WT, NOPS>(wp + (size_t)j * wstride, w0);
        load_w<WT, NOPS>(wp + (size_t)(j + 1) * wstride, w1);
        UT v0[FPT], v1[FPT];
        load_pack<UT, FPT>(u + (size_t)n0 * F + fofs, v0);
        load_pack<UT, FPT>(u + (size_t)n1 * F + fofs, v1);
#pragma unroll
        for (int o = 0; o < NOPS; o++) {
            AT a0 = (AT)w0[o];
#pragma unroll
            for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a0, (AT)v0[f], acc[o][f]);
        }
#pragma unroll
        for (int o = 0; o < NOPS; o++) {
            AT a1 = (AT)w1[o];
#pragma unroll
            for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a1, (AT)v1[f], acc[o][f]);
        }
    }
    if (j < width) {
        int n0 = ip[(size_t)j * n_pad];
        WT w0[NOPS];
        load_w<WT, NOPS>(wp + (size_t)j * wstride, w0);
        UT v0[FPT];
        load_pack<UT, FPT>(u + (size_t)n0 * F + fofs, v0);
#pragma unroll
        for (int o = 0; o < NOPS; o++) {
            AT a0 = (AT)w0[o];
#pragma unroll
            for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a0, (AT)v0[f], acc[o][f]);
        }
    }
}

// Maps a hardware block index to a logical tile so that each of the 8 XCDs (which receive
// blocks round-robin) works on one contiguous eighth of the rows and keeps its gathered
// neighbourhood in its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    int per = nb >> 3;  // nb is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

// Launch shape chosen for a fused kernel: FPT fields per thread, block threads.
struct EllTune {
    int fpt = 1;
    int block = 256;
};

int rbfx_ell_pick(rbfx_ell* ell, int n_fields, int field_type, int accum, int kind, EllTune* t);
