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
//
// Slots are processed in batches of UB: the UB neighbour records of a batch are gathered
// back to back, and the indices and weights of the next batch are requested before the
// current batch is consumed, so one memory round trip covers UB slots instead of the
// index-load -> gather dependency costing two round trips per slot. The summation order per
// accumulator stays j = 0, 1, ..., width-1 whatever UB is.
template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB>
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
    const UT* ub = u + fofs;
    int nb[UB];
    WT w[UB][NOPS];
#pragma unroll
    for (int q = 0; q < UB; q++) {
        if (q < width) {
            nb[q] = __builtin_nontemporal_load(ip + (size_t)q * n_pad);
            load_w<WT, NOPS>(wp + (size_t)q * wstride, w[q]);
        } else {
            nb[q] = 0;
#pragma unroll
            for (int o = 0; o < NOPS; o++) w[q][o] = (WT)0;
        }
    }
    for (int j0 = 0; j0 < width; j0 += UB) {
        UT v[UB][FPT];
#pragma unroll
        for (int q = 0; q < UB; q++) {
            if (j0 + q < width) load_pack<UT, FPT>(ub + (size_t)((unsigned)nb[q] * (unsigned)F), v[q]);
        }
        int nb2[UB];
        WT w2[UB][NOPS];
#pragma unroll
        for (int q = 0; q < UB; q++) {
            int j = j0 + UB + q;
            if (j < width) {
                nb2[q] = __builtin_nontemporal_load(ip + (size_t)j * n_pad);
                load_w<WT, NOPS>(wp + (size_t)j * wstride, w2[q]);
            } else {
                nb2[q] = 0;
#pragma unroll
                for (int o = 0; o < NOPS; o++) w2[q][o] = (WT)0;
            }
        }
#pragma unroll
        for (int q = 0; q < UB; q++) {
            if (j0 + q < width) {
#pragma unroll
                for (int o = 0; o < NOPS; o++) {
                    AT a = (AT)w[q][o];
#pragma unroll
                    for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a, (AT)v[q][f], acc[o][f]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < UB; q++) {
            nb[q] = nb2[q];
#pragma unroll
            for (int o = 0; o < NOPS; o++) w[q][o] = w2[q][o];
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
    int ub = 4;  // slots per gather batch (4 or 8)
};

int rbfx_ell_pick(rbfx_ell* ell, int n_fields, int field_type, int accum, int kind, EllTune* t);
