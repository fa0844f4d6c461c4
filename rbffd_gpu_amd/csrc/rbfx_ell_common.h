// Device layout of a fused ELL operator set and the gather-dot core shared by the SpMV,
// SpMM and RK4 stage kernels.
//
// Layout in HBM: rows in tiles of ELL_TILE (n_pad = n_rows rounded up to a whole tile),
// slot-major inside a tile. With e = (tile * width + j) * ELL_TILE + (row % ELL_TILE):
//   idx_t[e]                 int32: slot j of consecutive rows is one coalesced load
//   w_t[e * n_ops + o]       weights of the n_ops operators of one (slot,row) side by side,
//                            so one vector load fetches all of them; FP64 or FP32 storage
// All slots of a row lie within one tile's width * ELL_TILE entries (200 KB of FP32 weights
// at n = 50, 4 operators): a wave's loads over its whole stencil stay inside one or two
// pages, where a plain slot-major layout (slot stride = all rows) puts every slot on a page
// of its own and keeps the per-CU address-translation cache missing.
// Fields are node-major u[node * F + f], so the gathered neighbour record of FPT fields is
// one contiguous vector load. Pad rows carry index 0 and weight 0.
#pragma once
#include <type_traits>

#include "rbfx_internal.h"

#define ELL_TILE 256
#define ELL_ROW_PAD ELL_TILE

// Element offset of slot 0 of `row` in the tiled layout; slot j adds j * ELL_TILE.
__host__ __device__ __forceinline__ size_t ell_row_base(int row, int width) {
    return ((size_t)(row / ELL_TILE) * (size_t)width) * ELL_TILE + (size_t)(row % ELL_TILE);
}

struct rbfx_ell {
    rbfx_ctx* ctx = nullptr;
    int32_t n_rows = 0, n_cols = 0, width = 0, n_ops = 0, storage = RBFX_F64;
    int32_t n_pad = 0;
    uint32_t types[16] = {0};
    int32_t* idx_t = nullptr;  // device
    void* w_t = nullptr;       // device, double or float
    int32_t* order = nullptr;  // device, optional: row processed at position p (layout position p)
    int layout_version = 0;    // bumped whenever idx_t / w_t / order are replaced
    // LDS staging (optional): positions cut into n_ltiles tiles of at most lt rows (tile t =
    // positions loc_pos[t] .. loc_pos[t+1]); tile t gathers only from its distinct neighbour
    // nodes loc_nodes[loc_off[t] .. loc_off[t+1]) (at most k_max), which a block copies into
    // LDS once; loc_idx (same layout as idx_t) indexes that list.
    int lt = 0, k_max = 0, lds_limit = 0, n_ltiles = 0;
    int32_t* loc_pos = nullptr;
    int32_t* loc_off = nullptr;
    int32_t* loc_nodes = nullptr;
    uint16_t* loc_idx = nullptr;
};

// Row handled at processing position p (the identity unless a locality order is set).
__device__ __forceinline__ int ell_row_at(const int32_t* __restrict__ order, int p) { return order ? order[p] : p; }

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

// Streaming (non-temporal) forms of load_pack / store_pack for data touched once per kernel:
// 16-byte chunks with the nt hint when N * sizeof(T) is a multiple of 16, plain otherwise.
template <typename T, int N>
__device__ __forceinline__ void load_pack_nt(const T* __restrict__ p, T (&out)[N]) {
    constexpr int B = (int)(N * sizeof(T));
    if constexpr (B % 16 == 0) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        constexpr int C = 16 / sizeof(T);
#pragma unroll
        for (int c = 0; c < N / C; c++) {
            v4u q = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(p + c * C));
            Pack<T, C> r;
            __builtin_memcpy(&r, &q, 16);
#pragma unroll
            for (int i = 0; i < C; i++) out[c * C + i] = r.v[i];
        }
    } else {
        load_pack<T, N>(p, out);
    }
}
template <typename T, int N>
__device__ __forceinline__ void store_pack_nt(T* __restrict__ p, const T (&in)[N]) {
    constexpr int B = (int)(N * sizeof(T));
    if constexpr (B % 16 == 0) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        constexpr int C = 16 / sizeof(T);
#pragma unroll
        for (int c = 0; c < N / C; c++) {
            Pack<T, C> r;
#pragma unroll
            for (int i = 0; i < C; i++) r.v[i] = in[c * C + i];
            v4u q;
            __builtin_memcpy(&q, &r, 16);
            __builtin_nontemporal_store(q, reinterpret_cast<v4u*>(p + c * C));
        }
    } else {
        store_pack<T, N>(p, in);
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

// Weights of one (slot,row). FP32 weights carry the streaming (non-temporal) hint: they are
// read once per stage and should not displace the gathered field records from the caches
// (measured: 3 % faster stages). FP64 weights of multi-operator sets load normally (the hint
// measured 13-14 % slower stages); a single FP64 operator (the SpMV) takes the hint, which
// measured 9 % faster (124 -> 114 us at N = 1M, n = 50).
template <typename WT, int NOPS>
__device__ __forceinline__ void load_w_nt(const WT* __restrict__ p, WT (&w)[NOPS]) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    if constexpr (std::is_same<WT, float>::value && NOPS == 4) {
        f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else if constexpr (std::is_same<WT, float>::value) {
#pragma unroll
        for (int o = 0; o < NOPS; o++) w[o] = __builtin_nontemporal_load(p + o);
    } else if constexpr (NOPS == 1) {
        w[0] = __builtin_nontemporal_load(p);
    } else {
        load_w<WT, NOPS>(p, w);
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
                                               const UT* __restrict__ u, int row, int width, int F,
                                               int fofs, AT (&acc)[NOPS][FPT]) {
#pragma unroll
    for (int o = 0; o < NOPS; o++)
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[o][f] = (AT)0;
    const size_t base = ell_row_base(row, width);
    const int32_t* ip = idx_t + base;
    const WT* wp = w_t + base * NOPS;
    const size_t wstride = (size_t)ELL_TILE * NOPS;
    const UT* ub = u + fofs;
    int nb[UB];
    WT w[UB][NOPS];
#pragma unroll
    for (int q = 0; q < UB; q++) {
        if (q < width) {
            nb[q] = __builtin_nontemporal_load(ip + (size_t)q * ELL_TILE);
            load_w_nt<WT, NOPS>(wp +(size_t)q * wstride, w[q]);
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
                nb2[q] = __builtin_nontemporal_load(ip + (size_t)j * ELL_TILE);
                load_w_nt<WT, NOPS>(wp +(size_t)j * wstride, w2[q]);
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

// ---- shared slot loads across the G lanes of one row (G = 2 or 4) ----------------------
// When G lanes work on one row (G field groups), the plain core above has all G lanes load
// the same index and weights for every slot. Here lane q of the group loads the index and
// weights of slot j0 + q only (G slots per load instruction, no duplicates), and the values
// are broadcast inside the group with DPP quad permutes (VALU moves, no LDS, no memory).
// Fewer and narrower vector-memory instructions per slot: index loads drop by G, weight
// loads by G in count and in bytes.
template <int G, int S>
__device__ __forceinline__ int bcast_group(int v) {
    static_assert(G == 2 || G == 4, "DPP group broadcast handles groups of 2 or 4 lanes");
    // quad_perm selectors: every lane of the group reads lane S of the group
    constexpr int ctrl = (G == 4) ? (S * 0x55) : (S == 0 ? 0xA0 : 0xF5);
    // every lane of every row is written, so the destination's old value is never used
    return __builtin_amdgcn_mov_dpp(v, ctrl, 0xF, 0xF, false);
}
template <int G, int S>
__device__ __forceinline__ float bcast_group(float v) {
    return __int_as_float(bcast_group<G, S>(__float_as_int(v)));
}
template <int G, int S>
__device__ __forceinline__ double bcast_group(double v) {
    long long b = __double_as_longlong(v);
    int lo = bcast_group<G, S>((int)(b & 0xffffffffll));
    int hi = bcast_group<G, S>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int I, int N>
struct StaticFor {
    template <typename Fn>
    __device__ __forceinline__ static void run(Fn&& fn) {
        if constexpr (I < N) {
            fn(std::integral_constant<int, I>{});
            StaticFor<I + 1, N>::run(fn);
        }
    }
};

// Same result and summation order as ell_gather_dot; q = lane index inside the row's group.
// Requires UB % G == 0 and the G lanes of a group to be active together (they share a row).
template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int G>
__device__ __forceinline__ void ell_gather_dot_shared(const int32_t* __restrict__ idx_t, const WT* __restrict__ w_t,
                                                      const UT* __restrict__ u, int row, int width, int F,
                                                      int fofs, int q, AT (&acc)[NOPS][FPT]) {
    static_assert(UB % G == 0, "batch must be a multiple of the group size");
    constexpr int L = UB / G;  // own loads per batch
#pragma unroll
    for (int o = 0; o < NOPS; o++)
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[o][f] = (AT)0;
    const size_t base = ell_row_base(row, width);
    const int32_t* ip = idx_t + base;
    const WT* wp = w_t + base * NOPS;
    const size_t wstride = (size_t)ELL_TILE * NOPS;
    const UT* ub = u + fofs;
    int nbl[L];
    WT wl[L][NOPS];
#pragma unroll
    for (int l = 0; l < L; l++) {
        int j = l * G + q;
        if (j < width) {
            nbl[l] = __builtin_nontemporal_load(ip + (size_t)j * ELL_TILE);
            load_w_nt<WT, NOPS>(wp +(size_t)j * wstride, wl[l]);
        } else {
            nbl[l] = 0;
#pragma unroll
            for (int o = 0; o < NOPS; o++) wl[l][o] = (WT)0;
        }
    }
    for (int j0 = 0; j0 < width; j0 += UB) {
        // neighbour indices of the whole batch first: the gathers depend on them
        int nb[UB];
#pragma unroll
        for (int l = 0; l < L; l++)
            StaticFor<0, G>::run([&](auto S) { nb[l * G + decltype(S)::value] = bcast_group<G, decltype(S)::value>(nbl[l]); });
        UT v[UB][FPT];
#pragma unroll
        for (int k = 0; k < UB; k++)
            if (j0 + k < width) load_pack<UT, FPT>(ub + (size_t)((unsigned)nb[k] * (unsigned)F), v[k]);
        // next batch's own loads go to a second small buffer; this batch's weights stay in
        // wl and are broadcast only where the multiply-adds use them (few live registers)
        int nbn[L];
        WT wn[L][NOPS];
#pragma unroll
        for (int l = 0; l < L; l++) {
            int j = j0 + UB + l * G + q;
            if (j < width) {
                nbn[l] = __builtin_nontemporal_load(ip + (size_t)j * ELL_TILE);
                load_w_nt<WT, NOPS>(wp +(size_t)j * wstride, wn[l]);
            } else {
                nbn[l] = 0;
#pragma unroll
                for (int o = 0; o < NOPS; o++) wn[l][o] = (WT)0;
            }
        }
#pragma unroll
        for (int l = 0; l < L; l++) {
            StaticFor<0, G>::run([&](auto S) {
                constexpr int s = decltype(S)::value;
                const int k = l * G + s;
                if (j0 + k < width) {
#pragma unroll
                    for (int o = 0; o < NOPS; o++) {
                        AT a = (AT)bcast_group<G, s>(wl[l][o]);
#pragma unroll
                        for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a, (AT)v[k][f], acc[o][f]);
                    }
                }
            });
        }
#pragma unroll
        for (int l = 0; l < L; l++) {
            nbl[l] = nbn[l];
#pragma unroll
            for (int o = 0; o < NOPS; o++) wl[l][o] = wn[l][o];
        }
    }
}

// Shared loads with a deeper pipeline: the own index/weight loads run PD batches ahead (a
// ring of PD register slots, unrolled so every slot index is a compile-time constant), and
// the gathers of batch b+1 are issued before the multiply-adds of batch b wait for their
// own gathers. More bytes in flight per wave for the same occupancy; same summation order.
template <typename WT, int NOPS, int G, int L>
__device__ __forceinline__ void own_loads(const int32_t* __restrict__ ip, const WT* __restrict__ wp,
                                          size_t wstride, int width, int jb, int q, int (&nbl)[L],
                                          WT (&wl)[L][NOPS]) {
#pragma unroll
    for (int l = 0; l < L; l++) {
        int j = jb + l * G + q;
        if (j < width) {
            nbl[l] = __builtin_nontemporal_load(ip + (size_t)j * ELL_TILE);
            load_w_nt<WT, NOPS>(wp +(size_t)j * wstride, wl[l]);
        } else {
            nbl[l] = 0;
#pragma unroll
            for (int o = 0; o < NOPS; o++) wl[l][o] = (WT)0;
        }
    }
}

template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int G, int PD>
__device__ __forceinline__ void ell_gather_dot_pipe(const int32_t* __restrict__ idx_t, const WT* __restrict__ w_t,
                                                    const UT* __restrict__ u, int row, int width, int F,
                                                    int fofs, int q, AT (&acc)[NOPS][FPT]) {
    static_assert(UB % G == 0, "batch must be a multiple of the group size");
    constexpr int L = UB / G;
#pragma unroll
    for (int o = 0; o < NOPS; o++)
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[o][f] = (AT)0;
    const size_t base = ell_row_base(row, width);
    const int32_t* ip = idx_t + base;
    const WT* wp = w_t + base * NOPS;
    const size_t wstride = (size_t)ELL_TILE * NOPS;
    const UT* ub = u + fofs;
    const int nbatch = (width + UB - 1) / UB;
    int nbl[PD][L];
    WT wl[PD][L][NOPS];
    StaticFor<0, PD>::run([&](auto S) {
        constexpr int s = decltype(S)::value;
        own_loads<WT, NOPS, G, L>(ip, wp, wstride, width, s * UB, q, nbl[s], wl[s]);
    });
    UT vc[UB][FPT];
    auto gather = [&](const int (&nblr)[L], int jb, UT (&v)[UB][FPT]) {
#pragma unroll
        for (int l = 0; l < L; l++)
            StaticFor<0, G>::run([&](auto S) {
                constexpr int s = decltype(S)::value;
                const int k = l * G + s;
                const int nb = bcast_group<G, s>(nblr[l]);
                if (jb + k < width) load_pack<UT, FPT>(ub + (size_t)((unsigned)nb * (unsigned)F), v[k]);
            });
    };
    gather(nbl[0], 0, vc);
    for (int b0 = 0; b0 < nbatch; b0 += PD) {
        StaticFor<0, PD>::run([&](auto S) {
            constexpr int s = decltype(S)::value;
            constexpr int sn = (s + 1) % PD;
            const int b = b0 + s;
            if (b < nbatch) {
                const int jb = b * UB;
                UT vn[UB][FPT];
                if (b + 1 < nbatch) gather(nbl[sn], jb + UB, vn);
#pragma unroll
                for (int l = 0; l < L; l++)
                    StaticFor<0, G>::run([&](auto S2) {
                        constexpr int s2 = decltype(S2)::value;
                        const int k = l * G + s2;
                        if (jb + k < width) {
#pragma unroll
                            for (int o = 0; o < NOPS; o++) {
                                AT a = (AT)bcast_group<G, s2>(wl[s][l][o]);
#pragma unroll
                                for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a, (AT)vc[k][f], acc[o][f]);
                            }
                        }
                    });
                own_loads<WT, NOPS, G, L>(ip, wp, wstride, width, jb + PD * UB, q, nbl[s], wl[s]);
#pragma unroll
                for (int k = 0; k < UB; k++)
#pragma unroll
                    for (int f = 0; f < FPT; f++) vc[k][f] = vn[k][f];
            }
        });
    }
}

// ---- LDS-staged gathers ---------------------------------------------------------------
// A block owns one tile of lt consecutive positions. It first copies the tile's distinct
// neighbour records (F values each; F * sizeof(UT) a multiple of 16 bytes) from the field
// array into LDS, 16 bytes per thread per step, so each record crosses the memory system
// once per tile instead of once per (row, slot); the rows then gather from LDS through
// 16-bit local indices.
template <typename UT>
__device__ __forceinline__ void ell_stage_tile(const UT* __restrict__ u, int F, const int32_t* __restrict__ nodes, int K,
                                               UT* lds) {
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const int C = F * (int)sizeof(UT) / 16;
    const int total = K * C, B = blockDim.x;
    // Rounds of up to SI chunks per thread: all node ids of a round are requested, then all
    // records, then the LDS writes, so a round costs two memory latencies however many chunks
    // it holds (a plain loop waits twice per chunk).
    constexpr int SI = 8;
    for (int q0 = threadIdx.x; q0 < total; q0 += SI * B) {
        int kk[SI];
        v4u rec[SI];
#pragma unroll
        for (int i = 0; i < SI; i++) {
            const int q = q0 + i * B;
            kk[i] = q < total ? nodes[q / C] : 0;
        }
#pragma unroll
        for (int i = 0; i < SI; i++) {
            const int q = q0 + i * B;
            if (q < total) rec[i] = reinterpret_cast<const v4u*>(u + (size_t)kk[i] * F)[q - (q / C) * C];
        }
#pragma unroll
        for (int i = 0; i < SI; i++) {
            const int q = q0 + i * B;
            if (q < total) reinterpret_cast<v4u*>(lds)[q] = rec[i];
        }
    }
}

// Branch-free own loads: a slot past the stencil end loads slot width-1 again and gets weight
// 0, so the batch runs unguarded; its multiply-adds add 0 * (finite value) and leave the
// accumulators bit for bit unchanged.
template <int G, int L, int NOPS, typename WT>
__device__ __forceinline__ void own_loads_loc(const uint16_t* __restrict__ lip, const WT* __restrict__ wp,
                                              size_t wstride, int width, int jb, int q, int (&nbl)[L],
                                              WT (&wl)[L][NOPS]) {
#pragma unroll
    for (int l = 0; l < L; l++) {
        const int j = jb + l * G + q;
        const int jc = j < width ? j : width - 1;
        nbl[l] = (int)__builtin_nontemporal_load(lip + (size_t)jc * ELL_TILE);
        WT w[NOPS];
        load_w_nt<WT, NOPS>(wp +(size_t)jc * wstride, w);
#pragma unroll
        for (int o = 0; o < NOPS; o++) wl[l][o] = j < width ? w[o] : (WT)0;
    }
}

// Same result and summation order as ell_gather_dot, reading the neighbour records from the
// staged LDS copy: lane q of the row's G-lane group loads the local index and weights of
// slot jb + q, PD batches ahead, and the values are broadcast in the group (DPP). Records
// are F = FPT * G values apart in LDS.
template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int G, int PD>
__device__ __forceinline__ void ell_gather_dot_lds(const uint16_t* __restrict__ loc_idx, const WT* __restrict__ w_t,
                                                   const UT* lds, int p, int width, int fofs, int q,
                                                   AT (&acc)[NOPS][FPT]) {
    static_assert(UB % G == 0, "batch must be a multiple of the group size");
    constexpr int L = UB / G;
    constexpr int F = FPT * G;
#pragma unroll
    for (int o = 0; o < NOPS; o++)
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[o][f] = (AT)0;
    const size_t base = ell_row_base(p, width);
    const uint16_t* lip = loc_idx + base;
    const WT* wp = w_t + base * NOPS;
    const size_t wstride = (size_t)ELL_TILE * NOPS;
    const UT* lb = lds + fofs;
    const int nbatch = (width + UB - 1) / UB;
    int nbl[PD][L];
    WT wl[PD][L][NOPS];
    StaticFor<0, PD>::run([&](auto S) {
        constexpr int s = decltype(S)::value;
        own_loads_loc<G, L, NOPS, WT>(lip, wp, wstride, width, s * UB, q, nbl[s], wl[s]);
    });
    for (int b0 = 0; b0 < nbatch; b0 += PD) {
        StaticFor<0, PD>::run([&](auto S) {
            constexpr int s = decltype(S)::value;
            const int b = b0 + s;
            if (b < nbatch) {
                const int jb = b * UB;
                UT v[UB][FPT];
#pragma unroll
                for (int l = 0; l < L; l++)
                    StaticFor<0, G>::run([&](auto S2) {
                        constexpr int s2 = decltype(S2)::value;
                        const int li = bcast_group<G, s2>(nbl[s][l]);
                        load_pack<UT, FPT>(lb + li * F, v[l * G + s2]);
                    });
#pragma unroll
                for (int l = 0; l < L; l++)
                    StaticFor<0, G>::run([&](auto S2) {
                        constexpr int s2 = decltype(S2)::value;
                        const int k = l * G + s2;
#pragma unroll
                        for (int o = 0; o < NOPS; o++) {
                            AT a = (AT)bcast_group<G, s2>(wl[s][l][o]);
#pragma unroll
                            for (int f = 0; f < FPT; f++) acc[o][f] = fma_acc<AT>(a, (AT)v[k][f], acc[o][f]);
                        }
                    });
                own_loads_loc<G, L, NOPS, WT>(lip, wp, wstride, width, jb + PD * UB, q, nbl[s], wl[s]);
            }
        });
    }
}

// FP32 variant of ell_gather_dot_lds for 16 fields on 4 lanes per row (FPT = 4, G = 4) whose
// multiply-adds run on the matrix cores: one v_mfma_f32_4x4x1_16b_f32 treats the 16 rows of a
// wave as 16 independent 4x4 blocks and adds, per block, the outer product of a 4-vector held
// one value per lane of the row's group (A) with another (B). Per slot k and field offset g:
//   A at lane q = w_q[row][k]                     (operator q)
//   B at lane q = u[idx_k][4q + g]                (the lane's own 16-byte record fragment)
//   D_g[o] at lane q += A(lane o) * B(lane q)     = acc[o][g] of the lane, field 4q + g.
// Each lane loads its operator's weight of every slot (the 4 lanes of a row read 16 adjacent
// bytes), so no weight broadcast is needed; the local node ids are still loaded once per
// group and broadcast (DPP). The k order per accumulator is the same as in the VALU cores.
// MFMA reads operands across the lanes of a block only, and a row's 4 lanes are active or
// inactive together, so rows that exited early do not disturb the others.
//
// The core is split so that a kernel can issue the first PD batches of index and weight loads
// (which do not read LDS) before its staging barrier: start() issues them, finish() runs the
// pipelined loop over the staged records.
template <int NOPS, int UB, int PD>
struct LdsMfmaGather {
    typedef float f4 __attribute__((ext_vector_type(4)));
    static constexpr int G = 4, L = UB / G, F = 16;
    static_assert(UB % G == 0 && NOPS <= 4, "batch of whole groups, at most 4 operators");
    const uint16_t* lip;
    const float* wq;
    size_t wstride;
    int width, q;
    int nbl[PD][L];
    float wl[PD][UB];

    // Loads of the batch starting at slot jb. Slots past the row's width read the clamped last
    // slot (in range); the consumer masks them. No branch and no use of the loaded values
    // here, so PD batches of loads stay in flight.
    __device__ __forceinline__ void own(int jb, int (&nb)[L], float (&w)[UB]) const {
#pragma unroll
        for (int l = 0; l < L; l++) {
            const int j = jb + l * G + q;
            nb[l] = (int)__builtin_nontemporal_load(lip + (size_t)min(j, width - 1) * ELL_TILE);
        }
#pragma unroll
        for (int k = 0; k < UB; k++) w[k] = __builtin_nontemporal_load(wq + (size_t)min(jb + k, width - 1) * wstride);
    }

    // Row p (processing position) of the operator set; q = lane index inside the row's group.
    __device__ __forceinline__ void start(const uint16_t* __restrict__ loc_idx, const float* __restrict__ w_t, int p,
                                          int width_, int q_) {
        width = width_;
        q = q_;
        const size_t base = ell_row_base(p, width);
        lip = loc_idx + base;
        // lanes q >= NOPS load operator 0's weights: their A values only reach row q of each
        // 4x4 block, i.e. the accumulators of operator q, which are not returned
        wq = w_t + base * NOPS + (q < NOPS ? q : 0);
        wstride = (size_t)ELL_TILE * NOPS;
        StaticFor<0, PD>::run([&](auto S) { own(decltype(S)::value * UB, nbl[decltype(S)::value], wl[decltype(S)::value]); });
    }

    __device__ __forceinline__ void finish(const float* lds, int fofs, float (&acc)[NOPS][4]) {
        f4 D[4];
#pragma unroll
        for (int g = 0; g < 4; g++) D[g] = f4{0.f, 0.f, 0.f, 0.f};
        const float* lb = lds + fofs;
        const int nbatch = (width + UB - 1) / UB;
        // MASK: slots at or past the width get weight 0 (0 * a finite record: the accumulators
        // do not change)
        auto consume = [&](int (&nb)[L], float (&w)[UB], bool mask, int jb) {
            float v[UB][4];
#pragma unroll
            for (int l = 0; l < L; l++)
                StaticFor<0, G>::run([&](auto S2) {
                    constexpr int s2 = decltype(S2)::value;
                    const int li = bcast_group<G, s2>(nb[l]);
                    load_pack<float, 4>(lb + li * F, v[l * G + s2]);
                });
#pragma unroll
            for (int k = 0; k < UB; k++) {
                const float a = mask ? w[k] * (jb + k < width ? 1.f : 0.f) : w[k];
#pragma unroll
                for (int g = 0; g < 4; g++) D[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, v[k][g], D[g], 0, 0, 0);
            }
        };
        // the pipelined loop takes whole batches only, a multiple of PD of them; the rest (at
        // most PD - 1 whole batches and the partial one) are already loaded when it ends
        const int nfull = (width / UB) / PD * PD;
        for (int b0 = 0; b0 < nfull; b0 += PD) {
            StaticFor<0, PD>::run([&](auto S) {
                constexpr int s = decltype(S)::value;
                consume(nbl[s], wl[s], false, 0);
                // keep the reloads after the last use of the registers they overwrite: hoisted
                // above it they need fresh registers and loop-end copies that wait for every load
                __builtin_amdgcn_sched_barrier(0);
                own((b0 + s + PD) * UB, nbl[s], wl[s]);
            });
        }
        StaticFor<0, PD>::run([&](auto S) {
            constexpr int s = decltype(S)::value;
            if (nfull + s < nbatch) consume(nbl[s], wl[s], true, (nfull + s) * UB);
        });
#pragma unroll
        for (int o = 0; o < NOPS; o++)
#pragma unroll
            for (int g = 0; g < 4; g++) acc[o][g] = D[g][o];
    }
};

template <int NOPS, int UB, int PD>
__device__ __forceinline__ void ell_gather_dot_lds_mfma(const uint16_t* __restrict__ loc_idx,
                                                        const float* __restrict__ w_t, const float* lds, int p,
                                                        int width, int fofs, int q, float (&acc)[NOPS][4]) {
    LdsMfmaGather<NOPS, UB, PD> g;
    g.start(loc_idx, w_t, p, width, q);
    g.finish(lds, fofs, acc);
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
    int ub = 4;     // slots per gather batch (4 or 8)
    int share = 0;  // 1: the G = F/fpt lanes of a row split the index/weight loads (G = 2 or 4)
};

int rbfx_ell_pick(rbfx_ell* ell, int n_fields, int field_type, int accum, int kind, EllTune* t);
