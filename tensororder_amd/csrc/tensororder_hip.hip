// MI355X (gfx950) contraction executor: kernels and the C ABI of include/tensororder_hip.h.
//
// Replaces the arithmetic of the reference's execution path: leaf materialisation per
// slice (tensor_network_constructions.py:69-99, :144-152 via tensor.py:66-68), numpy
// tensordot on every join of the post-order walk (tensor_network.pyx:142-156,
// numpy_apis.py:42-43) and the FP64 sum of slice scalars (base_api.py:17-28).
//
// Design (CDNA4, wave64, 256-thread blocks):
//  * every tensor is an all-dimension-2 array, so a join is an index-bit gather: the
//    kernels never permute operands, they place bits with pdep tables instead;
//  * the streamed operand X is read once, coalesced along its free bits; the other
//    operand's slab (y chunk x 2^s) is staged in LDS and read as broadcasts;
//  * small-output / large-shared joins split the shared range over blocks and finish
//    with a deterministic fixed-order sum;
//  * a batch of slices runs in one launch per join (batch stride 0 = shared tensor);
//  * FP32 mode keeps one power-of-two exponent per tensor and batch item; consumers
//    rescale on load, producers fold max |entry| into an atomicMax on its float bits.
#include <hip/hip_runtime.h>
#include <mutex>
#include <type_traits>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "tensororder_hip.h"

#define TO_ABI_VERSION 6

namespace {

constexpr int kBlock = 256;

thread_local char g_err[512] = "";

void set_err(const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
}

int hip_fail(hipError_t e, const char* where) {
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TO_ERR_OOM : TO_ERR_HIP;
}

// ---------------------------------------------------------------------------------
// device helpers

// Device addresses arrive as integers in the descriptors; casting them to global-memory
// pointers (address space 1) makes the compiler emit global_* instead of flat_* memory
// instructions, so LDS waits do not also wait for outstanding HBM loads.
#define TO_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ TO_GLOBAL T* gptr(uint64_t p) {
    return (TO_GLOBAL T*)p;
}

__device__ __forceinline__ int64_t pdep(uint64_t v, const int8_t* pos, int n) {
    // The position bytes of 8 bits at a time are loaded before any is used (indices
    // clamped to n - 1, so every load is in bounds and none waits behind a loop exit):
    // one memory wait per 8 bits rather than one per bit.
    int64_t r = 0;
    for (int g = 0; g < n; g += 8) {
        int p[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) p[j] = pos[g + j < n ? g + j : n - 1];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (g + j < n) r |= (int64_t)((v >> (g + j)) & 1ull) << p[j];
    }
    return r;
}

// Bit positions spread across the lanes of a wave: lane i < n holds 1 << pos[i], loaded
// with one memory access per lane instead of one dependent access per bit.  Made where
// every lane of the wave runs (v_readlane reads a lane whether or not it is active
// later); a deposit then costs a few ALU instructions per bit and no memory wait.
struct LaneBits {
    uint32_t lo, hi;
};

__device__ __forceinline__ LaneBits lane_bits(const int8_t* pos, int n) {
    const int lane = threadIdx.x & 63;
    const uint64_t b = lane < n ? 1ull << pos[lane] : 0ull;
    return LaneBits{(uint32_t)b, (uint32_t)(b >> 32)};
}

// pdep(v, pos + first, n) from lane-held positions (first + n <= 64)
__device__ __forceinline__ int64_t pdep_l(uint64_t v, LaneBits p, int first, int n) {
    uint32_t lo = 0, hi = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t m = 0u - (uint32_t)((v >> i) & 1ull);
        lo |= m & (uint32_t)__builtin_amdgcn_readlane((int)p.lo, first + i);
        hi |= m & (uint32_t)__builtin_amdgcn_readlane((int)p.hi, first + i);
    }
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// a value known to be the same in every lane, moved to scalar registers
__device__ __forceinline__ int64_t uniform64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// index of the descriptor whose block range holds block bid: read from the launch's
// block -> descriptor map when it has one (one load), else by binary search over block0
// (about log2(n) dependent loads); the result is the same for every lane of the block, and
// is marked wave-uniform so descriptor fields load as scalars
template <class D>
__device__ __forceinline__ int find_desc(const D* __restrict__ descs, int n, int bid,
                                         const int32_t* __restrict__ bmap) {
    if (bmap) return __builtin_amdgcn_readfirstlane(bmap[bid]);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (descs[mid].block0 <= bid) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// frexp exponent of a non-negative float given by its bits (0 for zero).
__device__ __forceinline__ int scale_exp(uint32_t bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f)) return 0;
    int e;
    frexpf(m, &e);
    return e;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    for (int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBlock / 64; ++w) r += red[w];
    }
    return r;  // valid in thread 0
}

// Operand scale factor 2^-e for batch item b, and the exponent it contributes.
struct OpScale {
    float f;
    int64_t e;
};

__device__ __forceinline__ OpScale op_scale(uint64_t recs, int64_t stride, int64_t b) {
    const TO_GLOBAL to_scale_rec* r = gptr<const to_scale_rec>(recs) + b * stride;
    const int e = scale_exp(r->maxbits);
    return OpScale{ldexpf(1.0f, -e), r->exp + e};
}

// Raise the running max |entry| of batch item b to m (non-negative floats compare as their
// bits).  The atomic is skipped when the stored max (read from L2, so never above the true
// current value) is already at least m: all the waves of a launch publish into one record,
// and atomics on one address are serialised, so most of them would otherwise queue.
__device__ __forceinline__ void note_max(uint64_t recs, int64_t b, float m) {
    TO_GLOBAL to_scale_rec* r = gptr<to_scale_rec>(recs) + b;
    uint32_t* p = (uint32_t*)&r->maxbits;
    const uint32_t bits = __float_as_uint(m);
    if (bits > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, bits);
}

__device__ __forceinline__ void note_exp(uint64_t recs, int64_t b, int64_t e) {
    TO_GLOBAL to_scale_rec* r = gptr<to_scale_rec>(recs) + b;
    r->exp = e;
}

template <typename T>
__device__ __forceinline__ T absval(T v) {
    return v < T(0) ? -v : v;
}

// The general streaming joins accumulate float operands in double: the products of two
// floats are exact in double and the output is rounded once, when it is stored.  These
// kernels are bound by HBM bytes, so the double arithmetic costs nothing measurable.
template <typename T>
using Acc = double;

// ---------------------------------------------------------------------------------
// leaf gather: per batch item, copy row sum_g coef[g]*bit_g(slice) of each leaf table

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_leaf_gather(const to_leaf_desc* __restrict__ descs,
                                                        int ndesc, const int32_t* __restrict__ bmap,
                                                        const int64_t* __restrict__ wave) {
    const to_leaf_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int rank = d->rank;
    const int64_t E = 1ll << rank;
    const int64_t total = (int64_t)d->nbatch << rank;
    const int64_t begin = wave ? wave[0] : 0;
    const int G = d->ngroups;
    const TO_GLOBAL int64_t* coef = gptr<const int64_t>(d->coef);
    const TO_GLOBAL T* src = gptr<const T>(d->src);
    TO_GLOBAL T* dst = gptr<T>(d->dst);
    const int64_t stride = (int64_t)d->nblocks * kBlock;
    for (int64_t i = (int64_t)(blockIdx.x - d->block0) * kBlock + threadIdx.x; i < total; i += stride) {
        const int64_t b = i >> rank;
        const int64_t e = i & (E - 1);
        const int64_t sidx = begin + b;
        int64_t row = 0;
        for (int g = 0; g < G; ++g) row += coef[g] * ((sidx >> (G - 1 - g)) & 1);
        dst[i] = src[row * E + e];
        if (SCALED && e == 0 && d->rec) {
            TO_GLOBAL to_scale_rec* r = gptr<to_scale_rec>(d->rec) + b;
            r->exp = d->exp0;
            r->maxbits = __float_as_uint(1.0f);
        }
    }
}

// ---------------------------------------------------------------------------------
// leaf build: OR and variable leaf tables written from their descriptions, once per plan

template <typename T>
__global__ __launch_bounds__(kBlock) void k_leaf_build(const to_leaf_build_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    const to_leaf_build_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int rank = d->rank;
    const int64_t E = 1ll << rank;
    const int64_t all = E - 1;
    TO_GLOBAL T* dst = gptr<T>(d->dst);
    const int64_t stride = (int64_t)d->nblocks * kBlock;
    for (int64_t e = (int64_t)(blockIdx.x - d->block0) * kBlock + threadIdx.x; e < E; e += stride) {
        const int64_t i = pdep(e, d->pos, rank);  // the leaf's own C-order index
        bool one;
        if (d->kind == TO_LEAF_VAR) {
            one = rank > 0 && i == all;
            const double v = one ? d->val1 : ((rank == 0 || i == 0) ? d->val0 : 0.0);
            dst[e] = (T)v;
            continue;
        }
        const int64_t l = i ^ d->negmask;
        if (d->out_bit < 0) {
            one = l != 0;
        } else {
            const int64_t ob = 1ll << d->out_bit;
            one = ((l & ob) != 0) == ((l & ~ob) != 0);
        }
        dst[e] = (T)(one ? d->val1 : d->val0);
    }
}

// ---------------------------------------------------------------------------------
// streaming join: each thread owns one free index f of X and YC free indices y of Y.
// LDS: kx[2^s], ky[2^s] (uint32 offsets of the shared bits), then (fx >= 8 only) the
// Y slab ylds[YC][2^s] of the current (batch item, y chunk), held as doubles so the inner
// loop converts only the X values.

template <typename T, bool SCALED, int YC>
__device__ __forceinline__ void join_stream_body(const to_join_desc* __restrict__ d, const int local,
                                                 unsigned char* smem) {
    const int tid = threadIdx.x;
    const int fx = d->fx, fy = d->fy, s = d->s;
    const int K = 1 << s;
    const int lnych = d->lnych;
    const int nych = 1 << lnych;
    const int nbatch = d->nbatch;
    const TO_GLOBAL T* __restrict__ X = gptr<const T>(d->x);
    const TO_GLOBAL T* __restrict__ Y = gptr<const T>(d->y);
    TO_GLOBAL T* __restrict__ C = gptr<T>(d->c);
    const int64_t xb = d->xb, yb = d->yb, cb = d->cb;
    const int shf = d->shf, shy = d->shy;
    const int NY = 1 << fy;

    uint32_t* kx = reinterpret_cast<uint32_t*>(smem);
    uint32_t* ky = kx + K;
    Acc<T>* ylds = reinterpret_cast<Acc<T>*>(ky + K);  // held as the accumulator type
    const LaneBits XF = lane_bits(d->xf_pos, fx), XS = lane_bits(d->xs_pos, s);
    const LaneBits YF = lane_bits(d->yf_pos, fy), YS = lane_bits(d->ys_pos, s);
    for (int k = tid; k < K; k += kBlock) {
        kx[k] = (uint32_t)pdep_l(k, XS, 0, s);
        ky[k] = (uint32_t)pdep_l(k, YS, 0, s);
    }
    __syncthreads();

    
    const int nblk = d->nblocks;

    if (fx >= 8) {
        // tiles of 256 consecutive f inside one batch item, y chunks innermost so a block
        // revisits the same X rows while they are in cache.  Tile t = ((batch * nft + ft)
        // << lnych) | ych; the counters advance incrementally (no division per tile).
        const int hb = fx - 8;
        const int64_t nft = 1ll << hb;
        const int64_t ntiles = (int64_t)nbatch << (hb + lnych);
        const int64_t t0 = ntiles * local / nblk, t1 = ntiles * (local + 1) / nblk;
        const int64_t xlo = pdep_l(tid, XF, 0, 8);
        // free-bit positions increase with the bit index, so the offsets of consecutive ft
        // follow x' = ((x | ~mask) + 1) & mask
        const int64_t hmask = pdep_l(~0ull, XF, 8, hb);
        int ych = (int)(t0 & (nych - 1));
        int64_t ft = (t0 >> lnych) & (nft - 1);
        int64_t batch = t0 >> (lnych + hb);
        int64_t xhi = pdep_l(ft, XF, 8, hb);
        int64_t staged = -1;
        // Y-free bits below the f bits, all of them in one chunk: stage the outputs
        const bool stage_out = (d->stage & 1) != 0;
        const bool stage_in = (d->stage & 2) != 0;
        T* otile = reinterpret_cast<T*>(ylds + YC * K);
        T* xtile = otile + (stage_out ? (kBlock << fy) : 0);
        // y bits lowest in C, C and its batch stride 16-byte aligned
        const bool vec_c = shy == 0 && ((d->c | (uint64_t)(cb * (int64_t)sizeof(T))) & 15) == 0;
        float mrun = 0.0f;
        for (int64_t t = t0; t < t1; ++t) {
            const int64_t key = batch * nych + ych;
            const int ny = min(YC, NY - ych * YC);
            float sx = 1.0f, sy = 1.0f;
            int64_t ex = 0, ey = 0;
            if (SCALED) {
                const OpScale a = op_scale(d->xs, d->xsb, batch);
                const OpScale c = op_scale(d->ys, d->ysb, batch);
                sx = a.f; ex = a.e; sy = c.f; ey = c.e;
            }
            if (key != staged) {
                __syncthreads();
                const auto Yb = Y + batch * yb;
                for (int i = tid; i < YC * K; i += kBlock) {
                    const int yc = i >> s, k = i & (K - 1);
                    T v = T(0);
                    if (yc < ny) v = Yb[pdep_l(ych * YC + yc, YF, 0, fy) + ky[k]] * T(sy);
                    ylds[i] = (Acc<T>)v;
                }
                __syncthreads();
                staged = key;
            }
            const int64_t f = (ft << 8) | tid;
            const auto Xp = X + batch * xb + xhi + xlo;
            if (stage_in) {
                // the tile's X entries are the contiguous run of 2^(8+s) entries at the
                // high-bit offset: load it linearly, then read this thread's entries in LDS
                const auto src = X + batch * xb + xhi;
                __syncthreads();
                for (int i = tid; i < (kBlock << s); i += kBlock) xtile[i] = src[i];
                __syncthreads();
            }
            Acc<T> acc[YC];
#pragma unroll
            for (int yc = 0; yc < YC; ++yc) acc[yc] = 0.0;
            for (int k0 = 0; k0 < K; k0 += 8) {
                T xv[8];
                if (stage_in) {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) xv[kk] = (k0 + kk < K) ? xtile[xlo + kx[k0 + kk]] * T(sx) : T(0);
                } else {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) xv[kk] = (k0 + kk < K) ? Xp[kx[k0 + kk]] * T(sx) : T(0);
                }
#pragma unroll
                for (int yc = 0; yc < YC; ++yc) {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) {
                        if (k0 + kk < K) acc[yc] += (Acc<T>)xv[kk] * (Acc<T>)ylds[yc * K + k0 + kk];
                    }
                }
            }
            auto Cb = C + batch * cb;
            float m = 0.0f;
            if (SCALED) {
#pragma unroll
                for (int yc = 0; yc < YC; ++yc)
                    if (yc < ny) m = fmaxf(m, (float)absval((T)acc[yc]));
            }
            if (stage_out) {
                // the tile's outputs are one contiguous run of 256 * 2^fy entries: stage
                // them in LDS and store them with consecutive lanes on consecutive entries
                __syncthreads();
#pragma unroll
                for (int yc = 0; yc < YC; ++yc)
                    if (yc < NY) otile[tid * NY + yc] = (T)acc[yc];
                __syncthreads();
                auto dst = Cb + (ft << (8 + fy));
                for (int i = tid; i < (kBlock << fy); i += kBlock) dst[i] = otile[i];
            } else if (YC * sizeof(T) >= 16 && vec_c && ny == YC) {
                // the chunk's y values are consecutive entries (y bits lowest): 16-byte
                // stores, aligned since YC * ych and 2^shf >= YC are multiples of the vector
                typedef T V16 __attribute__((ext_vector_type(16 / sizeof(T))));
                constexpr int VE = 16 / sizeof(T);
                auto dst = Cb + (f << shf) + (int64_t)ych * YC;
#pragma unroll
                for (int yc = 0; yc < YC; yc += VE) {
                    V16 w;
#pragma unroll
                    for (int e = 0; e < VE; ++e) w[e] = (T)acc[yc + e];
                    *reinterpret_cast<TO_GLOBAL V16*>(dst + yc) = w;
                }
            } else {
#pragma unroll
                for (int yc = 0; yc < YC; ++yc) {
                    if (yc < ny) {
                        const int64_t y = (int64_t)ych * YC + yc;
                        Cb[(f << shf) | (y << shy)] = (T)acc[yc];
                    }
                }
            }
            if (SCALED) {
                // the max of a batch item's outputs is kept per thread and published
                // (one atomic per wave) when the block moves to the next batch item
                mrun = fmaxf(mrun, m);
                if (tid == 0 && ft == 0 && ych == 0) note_exp(d->cs, batch, ex + ey);
                const bool last = (t + 1 == t1) || (ych + 1 == nych && ft + 1 == nft);
                if (last) {
                    const float w = wave_max(mrun);
                    if ((tid & 63) == 0) note_max(d->cs, batch, w);
                    mrun = 0.0f;
                }
            }
            if (++ych == nych) {
                ych = 0;
                if (++ft == nft) {
                    ft = 0;
                    xhi = 0;
                    ++batch;
                } else {
                    xhi = ((xhi | ~hmask) + 1) & hmask;
                }
            }
        }
    } else {
        // tiles of 256 consecutive (batch item, f) pairs; Y read through the cache
        const int fmask = (1 << fx) - 1;
        const int64_t nL = (int64_t)nbatch << fx;
        const int64_t nLt = (nL + kBlock - 1) / kBlock;
        const int64_t ntiles = nLt << lnych;
        const int64_t t0 = ntiles * local / nblk, t1 = ntiles * (local + 1) / nblk;
        const int64_t xoff = pdep_l(tid & fmask, XF, 0, fx);
        const int f = tid & fmask;
        const bool vec_c = shy == 0 && ((d->c | (uint64_t)(cb * (int64_t)sizeof(T))) & 15) == 0;
        int ych = (int)(t0 & (nych - 1));
        int64_t lt = t0 >> lnych;
        for (int64_t t = t0; t < t1; ++t) {
            const int64_t L = lt * kBlock + tid;
            float m = 0.0f;
            if (L < nL) {
                const int64_t batch = L >> fx;
                const int ny = min(YC, NY - ych * YC);
                float sx = 1.0f, sy = 1.0f;
                int64_t ex = 0, ey = 0;
                if (SCALED) {
                    const OpScale a = op_scale(d->xs, d->xsb, batch);
                    const OpScale c = op_scale(d->ys, d->ysb, batch);
                    sx = a.f; ex = a.e; sy = c.f; ey = c.e;
                }
                const auto Xp = X + batch * xb + xoff;
                const auto Yb = Y + batch * yb;
                int64_t yoff[YC];
#pragma unroll
                for (int yc = 0; yc < YC; ++yc) yoff[yc] = (yc < ny) ? pdep_l(ych * YC + yc, YF, 0, fy) : 0;
                Acc<T> acc[YC];
#pragma unroll
                for (int yc = 0; yc < YC; ++yc) acc[yc] = 0.0;
                for (int k = 0; k < K; ++k) {
                    const T xv = Xp[kx[k]] * T(sx);
#pragma unroll
                    for (int yc = 0; yc < YC; ++yc) {
                        if (yc < ny) acc[yc] += (Acc<T>)xv * (Acc<T>)(Yb[yoff[yc] + ky[k]] * T(sy));
                    }
                }
                auto Cb = C + batch * cb;
                if (YC * sizeof(T) >= 16 && vec_c && ny == YC) {
                    // consecutive y entries (y bits lowest): 16-byte stores
                    typedef T V16 __attribute__((ext_vector_type(16 / sizeof(T))));
                    constexpr int VE = 16 / sizeof(T);
                    auto dst = Cb + ((int64_t)f << shf) + (int64_t)ych * YC;
#pragma unroll
                    for (int yc = 0; yc < YC; yc += VE) {
                        V16 w;
#pragma unroll
                        for (int e = 0; e < VE; ++e) {
                            w[e] = (T)acc[yc + e];
                            if (SCALED) m = fmaxf(m, (float)absval(w[e]));
                        }
                        *reinterpret_cast<TO_GLOBAL V16*>(dst + yc) = w;
                    }
                } else {
#pragma unroll
                    for (int yc = 0; yc < YC; ++yc) {
                        if (yc < ny) {
                            const int64_t y = (int64_t)ych * YC + yc;
                            const T v = (T)acc[yc];
                            Cb[((int64_t)f << shf) | (y << shy)] = v;
                            if (SCALED) m = fmaxf(m, (float)absval(v));
                        }
                    }
                }
                if (SCALED && f == 0 && ych == 0) note_exp(d->cs, batch, ex + ey);
            }
            if (SCALED) {
                // lanes of one batch item are an aligned group of 2^fx lanes (or the whole
                // wave): reduce inside the group, its first lane publishes the max
                const int group = fx < 6 ? (1 << fx) : 64;
                for (int off = 1; off < group; off <<= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
                if (L < nL && (tid & (group - 1)) == 0) note_max(d->cs, L >> fx, m);
            }
            if (++ych == nych) {
                ych = 0;
                ++lt;
            }
        }
    }
}

template <typename T, bool SCALED, int YC>
__global__ __launch_bounds__(kBlock) void k_join_stream(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_stream_body<T, SCALED, YC>(d, blockIdx.x - d->block0, smem);
}

// ---------------------------------------------------------------------------------
// split-K join: block (o, sp) sums its share of the 2^s shared indices of output o.

template <typename T, bool SCALED>
__device__ __forceinline__ void join_splitk_body(const to_join_desc* __restrict__ d, const int64_t local) {
    __shared__ double red[kBlock / 64];
    const int tid = threadIdx.x;
    const int fx = d->fx, fy = d->fy, s = d->s;
    const int nsplit = d->nsplit;
    const int64_t o = local / nsplit;
    const int sp = (int)(local % nsplit);
    const int64_t batch = o >> (fx + fy);
    const int64_t f = (o >> fy) & ((1ll << fx) - 1);
    const int64_t y = o & ((1ll << fy) - 1);
    const TO_GLOBAL T* X = gptr<const T>(d->x) + batch * d->xb + pdep(f, d->xf_pos, fx);
    const TO_GLOBAL T* Y = gptr<const T>(d->y) + batch * d->yb + pdep(y, d->yf_pos, fy);
    float sx = 1.0f, sy = 1.0f;
    if (SCALED) {
        sx = op_scale(d->xs, d->xsb, batch).f;
        sy = op_scale(d->ys, d->ysb, batch).f;
    }
    double sum = 0.0;
    if (s >= 8) {
        const int64_t chunk_hi = (1ll << (s - 8)) / nsplit;  // 256-wide steps per split
        const int64_t xlo = pdep(tid, d->xs_pos, 8), ylo = pdep(tid, d->ys_pos, 8);
        // When the high shared bits of both operands lie in ascending positions (always so
        // after a permute to [free][shared] form), pdep(kh + 1) is pdep(kh) plus one inside
        // the bit mask (carries skip the bits outside it), so the step needs no pdep loop
        bool ascending = true;
        for (int i = 9; i < s; ++i)
            ascending = ascending && d->xs_pos[i - 1] < d->xs_pos[i] && d->ys_pos[i - 1] < d->ys_pos[i];
        if (ascending) {
            const int64_t xmask = uniform64(pdep(~0ull, d->xs_pos + 8, s - 8));
            const int64_t ymask = uniform64(pdep(~0ull, d->ys_pos + 8, s - 8));
            int64_t xh = uniform64(pdep(sp * chunk_hi, d->xs_pos + 8, s - 8));
            int64_t yh = uniform64(pdep(sp * chunk_hi, d->ys_pos + 8, s - 8));
#pragma unroll 4
            for (int64_t kh = 0; kh < chunk_hi; ++kh) {
                sum += (double)(X[xh + xlo] * T(sx)) * (double)(Y[yh + ylo] * T(sy));
                xh = ((xh | ~xmask) + 1) & xmask;
                yh = ((yh | ~ymask) + 1) & ymask;
            }
        } else {
            for (int64_t kh = sp * chunk_hi; kh < (sp + 1) * chunk_hi; ++kh) {
                const int64_t xo = pdep(kh, d->xs_pos + 8, s - 8) + xlo;
                const int64_t yo = pdep(kh, d->ys_pos + 8, s - 8) + ylo;
                sum += (double)(X[xo] * T(sx)) * (double)(Y[yo] * T(sy));
            }
        }
    } else if (tid < (1 << s)) {
        sum = (double)(X[pdep(tid, d->xs_pos, s)] * T(sx)) * (double)(Y[pdep(tid, d->ys_pos, s)] * T(sy));
    }
    sum = block_sum(sum, red);
    if (tid == 0) gptr<double>(d->partial)[o * nsplit + sp] = sum;
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_splitk(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_splitk_body<T, SCALED>(d, blockIdx.x - d->block0);
}

template <typename T, bool SCALED>
__device__ __forceinline__ void join_finish_body(const to_join_desc* __restrict__ d, const int64_t local) {
    const int fx = d->fx, fy = d->fy;
    const int nsplit = d->nsplit;
    const int64_t nout = (int64_t)d->nbatch << (fx + fy);
    const TO_GLOBAL double* partial = gptr<const double>(d->partial);
    TO_GLOBAL T* C = gptr<T>(d->c);
    const int64_t stride = (int64_t)d->nblocks * kBlock;
    // every lane runs every pass (outputs past the end only idle), so a pass's max can be
    // reduced across the wave and published with one atomic where the wave's outputs share
    // a batch item (one per output queued on the record when a finish covers many outputs)
    for (int64_t o0 = local * kBlock; o0 < nout; o0 += stride) {
        const int64_t o = o0 + threadIdx.x;
        const bool valid = o < nout;
        const int64_t batch = (valid ? o : nout - 1) >> (fx + fy);
        float m = 0.0f;
        if (valid) {
            double acc = 0.0;
            for (int sp = 0; sp < nsplit; ++sp) acc += partial[o * nsplit + sp];
            const T sum = (T)acc;
            const int64_t f = (o >> fy) & ((1ll << fx) - 1);
            const int64_t y = o & ((1ll << fy) - 1);
            C[batch * d->cb + ((f << d->shf) | (y << d->shy))] = sum;
            m = (float)absval(sum);
            if (SCALED && f == 0 && y == 0) {
                const int64_t e = op_scale(d->xs, d->xsb, batch).e + op_scale(d->ys, d->ysb, batch).e;
                note_exp(d->cs, batch, e);
            }
        }
        if (SCALED) {
            const int64_t b0 = uniform64(batch);
            if (__all(batch == b0)) {
                const float w = wave_max(m);
                if ((threadIdx.x & 63) == 0) note_max(d->cs, b0, w);
            } else if (valid) {
                note_max(d->cs, batch, m);
            }
        }
    }
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_finish(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_finish_body<T, SCALED>(d, blockIdx.x - d->block0);
}

// Many partial sums per output (nsplit >= 32, e.g. a dot product of two large tensors):
// one block per output, its threads sum strided partials and reduce in LDS.
template <typename T, bool SCALED>
__device__ __forceinline__ void join_finish_wide_body(const to_join_desc* __restrict__ d, const int64_t local) {
    __shared__ double red[kBlock / 64];
    const int fx = d->fx, fy = d->fy;
    const int nsplit = d->nsplit;
    const int64_t nout = (int64_t)d->nbatch << (fx + fy);
    const TO_GLOBAL double* partial = gptr<const double>(d->partial);
    TO_GLOBAL T* C = gptr<T>(d->c);
    for (int64_t o = local; o < nout; o += d->nblocks) {
        double acc = 0.0;
        for (int sp = threadIdx.x; sp < nsplit; sp += kBlock) acc += partial[o * nsplit + sp];
        acc = block_sum(acc, red);
        if (threadIdx.x == 0) {
            const T sum = (T)acc;
            const int64_t batch = o >> (fx + fy);
            const int64_t f = (o >> fy) & ((1ll << fx) - 1);
            const int64_t y = o & ((1ll << fy) - 1);
            C[batch * d->cb + ((f << d->shf) | (y << d->shy))] = sum;
            if (SCALED) {
                note_max(d->cs, batch, (float)absval(sum));
                if (f == 0 && y == 0) {
                    const int64_t e = op_scale(d->xs, d->xsb, batch).e + op_scale(d->ys, d->ysb, batch).e;
                    note_exp(d->cs, batch, e);
                }
            }
        }
    }
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_finish_wide(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_finish_wide_body<T, SCALED>(d, blockIdx.x - d->block0);
}

// One join given by value (to_tensordot): the kernels above without a descriptor table.
template <typename T, bool SCALED, int YC>
__global__ __launch_bounds__(kBlock) void k_join_stream_one(const to_join_desc desc) {
    extern __shared__ __align__(16) unsigned char smem[];
    join_stream_body<T, SCALED, YC>(&desc, blockIdx.x, smem);
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_splitk_one(const to_join_desc desc) {
    join_splitk_body<T, SCALED>(&desc, blockIdx.x);
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_finish_one(const to_join_desc desc) {
    join_finish_body<T, SCALED>(&desc, blockIdx.x);
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_finish_wide_one(const to_join_desc desc) {
    join_finish_wide_body<T, SCALED>(&desc, blockIdx.x);
}

// ---------------------------------------------------------------------------------
// bookkeeping kernels

__global__ __launch_bounds__(kBlock) void k_scale_clear(to_scale_rec* recs, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        recs[i].exp = 0;
        recs[i].maxbits = 0u;
        recs[i].pad = 0u;
    }
}

__global__ void k_wave_advance(int64_t* wave, int64_t step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) wave[0] += step;
}

// Store each listed instance's root value into its accumulator, stamping the device clock.
template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_root_store(const to_root_desc* __restrict__ descs, int ndesc) {
    const int64_t now = (int64_t)wall_clock64();
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < ndesc; i += gridDim.x * kBlock) {
        const to_root_desc& d = descs[i];
        const double v = (double)*gptr<const T>(d.root);
        int e = 0;
        const double m = frexp(v, &e);
        TO_GLOBAL to_accum* acc = gptr<to_accum>(d.acc);
        acc->mant = m;
        acc->exp = m == 0.0 ? 0 : (int64_t)e + (SCALED ? gptr<const to_scale_rec>(d.rec)->exp : 0);
        acc->count = 1;
        acc->pad = now;
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_begin(to_accum* acc, int64_t n, int64_t* start) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        acc[i].mant = 0.0;
        acc[i].exp = 0;
        acc[i].count = 0;
        acc[i].pad = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && start) *start = (int64_t)wall_clock64();
}

__global__ void k_step_begin(to_accum* acc, int64_t* wave) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        acc->mant = 0.0;
        acc->exp = 0;
        acc->count = 0;
        acc->pad = 0;
        wave[0] = 0;
    }
}

// Sum the root scalars of the valid batch items (slice index < wave[1]) into acc as
// (mantissa, exponent): align every item to the largest exponent, add in a fixed order.
template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_root_reduce(const T* __restrict__ roots,
                                                        const to_scale_rec* __restrict__ recs,
                                                        to_accum* acc, const int64_t* __restrict__ wave,
                                                        int nbatch) {
    __shared__ int64_t emax_s[kBlock / 64];
    __shared__ double red[kBlock / 64];
    const int tid = threadIdx.x;
    const int64_t begin = wave[0], end = wave[1];
    int64_t nvalid = end - begin;
    nvalid = nvalid < 0 ? 0 : (nvalid > nbatch ? nbatch : nvalid);
    int64_t emax = INT64_MIN;
    for (int64_t b = tid; b < nvalid; b += kBlock) {
        const double v = (double)roots[b];
        if (v != 0.0) {
            int e;
            frexp(v, &e);
            const int64_t ee = e + (SCALED ? recs[b].exp : 0);
            emax = ee > emax ? ee : emax;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(emax, off, 64);
        emax = o > emax ? o : emax;
    }
    if ((tid & 63) == 0) emax_s[tid >> 6] = emax;
    __syncthreads();
    emax = emax_s[0];
    for (int w = 1; w < kBlock / 64; ++w) emax = emax_s[w] > emax ? emax_s[w] : emax;
    double part = 0.0;
    if (emax != INT64_MIN) {
        for (int64_t b = tid; b < nvalid; b += kBlock) {
            const double v = (double)roots[b];
            if (v != 0.0) {
                int e;
                const double m = frexp(v, &e);
                int64_t sh = e + (SCALED ? recs[b].exp : 0) - emax;
                sh = sh < -2000 ? -2000 : sh;
                part += ldexp(m, (int)sh);
            }
        }
    }
    const double total = block_sum(part, red);
    if (tid == 0) {
        double am = acc->mant;
        int64_t ae = acc->exp;
        if (emax != INT64_MIN && total != 0.0) {
            if (am == 0.0) {
                am = total;
                ae = emax;
            } else {
                const int64_t E = ae > emax ? ae : emax;
                int64_t d1 = ae - E, d2 = emax - E;
                d1 = d1 < -2000 ? -2000 : d1;
                d2 = d2 < -2000 ? -2000 : d2;
                am = ldexp(am, (int)d1) + ldexp(total, (int)d2);
                ae = E;
            }
            int e2;
            const double m2 = frexp(am, &e2);
            am = m2;
            ae += e2;
        }
        acc->mant = am;
        acc->exp = am == 0.0 ? 0 : ae;
        acc->count += nvalid;
    }
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

// ---------------------------------------------------------------------------------
// streaming join, 16-byte form: the common skinny join whose X free bits start at bit 0
// of X, whose output puts all 2^fy y values of one f next to each other (shy == 0,
// shf == fy) and whose Y slab fits one chunk.  Each thread owns VF consecutive f, so it
// loads 16-byte X chunks and stores its VF * 2^fy outputs as 16-byte chunks; a block walks
// tiles of 256 * VF consecutive f of one batch item.

template <typename T, bool SCALED, int YC>
__device__ __forceinline__ void join_stream_vec_body(const to_join_desc* __restrict__ d, const int local,
                                                     unsigned char* smem) {
    constexpr int VF = 16 / sizeof(T);
    constexpr int LVF = sizeof(T) == 4 ? 2 : 1;
    typedef typename Vec16<T>::type V;
    const int tid = threadIdx.x;
    const int fx = d->fx, fy = d->fy, s = d->s;
    const int K = 1 << s;
    const int nbatch = d->nbatch;
    const TO_GLOBAL T* __restrict__ X = gptr<const T>(d->x);
    const TO_GLOBAL T* __restrict__ Y = gptr<const T>(d->y);
    TO_GLOBAL T* __restrict__ C = gptr<T>(d->c);
    const int64_t xb = d->xb, yb = d->yb, cb = d->cb;

    uint32_t* kx = reinterpret_cast<uint32_t*>(smem);
    uint32_t* ky = kx + K;
    T* ylds = reinterpret_cast<T*>(smem + ((8 * K + 15) & ~15));
    const LaneBits XF = lane_bits(d->xf_pos, fx), XS = lane_bits(d->xs_pos, s);
    const LaneBits YF = lane_bits(d->yf_pos, fy), YS = lane_bits(d->ys_pos, s);
    for (int k = tid; k < K; k += kBlock) {
        kx[k] = (uint32_t)pdep_l(k, XS, 0, s);
        ky[k] = (uint32_t)pdep_l(k, YS, 0, s);
    }
    __syncthreads();

    const int nblk = d->nblocks;
    const int tb = 8 + LVF;  // f bits inside a tile
    const int hb = fx - tb;
    const int64_t nft = 1ll << hb;
    const int64_t ntiles = (int64_t)nbatch << hb;
    const int64_t t0 = ntiles * local / nblk, t1 = ntiles * (local + 1) / nblk;
    const int64_t xlo = pdep_l((uint64_t)tid << LVF, XF, 0, tb);
    const int64_t hmask = pdep_l(~0ull, XF, tb, hb);
    int64_t ft = t0 & (nft - 1);
    int64_t batch = t0 >> hb;
    int64_t xhi = pdep_l(ft, XF, tb, hb);
    int64_t staged = -1;
    float sx = 1.0f, sy = 1.0f;
    int64_t ex = 0, ey = 0;
    float mrun = 0.0f;
    for (int64_t t = t0; t < t1; ++t) {
        if (batch != staged) {
            if (SCALED) {
                const OpScale a = op_scale(d->xs, d->xsb, batch);
                const OpScale c = op_scale(d->ys, d->ysb, batch);
                sx = a.f; ex = a.e; sy = c.f; ey = c.e;
            }
            __syncthreads();
            const auto Yb = Y + batch * yb;
            for (int i = tid; i < YC * K; i += kBlock) {
                const int yc = i >> s, k = i & (K - 1);
                ylds[i] = Yb[pdep_l(yc, YF, 0, fy) + ky[k]] * T(sy);
            }
            __syncthreads();
            staged = batch;
        }
        const auto Xp = X + batch * xb + xhi + xlo;
        // float accumulators here: the 16-byte form keeps VF * YC sums live, which in
        // double would cost the occupancy this HBM-bound kernel needs
        T acc[VF][YC];
#pragma unroll
        for (int i = 0; i < VF; ++i)
#pragma unroll
            for (int y = 0; y < YC; ++y) acc[i][y] = T(0);
        for (int k0 = 0; k0 < K; k0 += 4) {
            V xv[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                if (k0 + kk < K) xv[kk] = *reinterpret_cast<const TO_GLOBAL V*>(Xp + kx[k0 + kk]) * T(sx);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                if (k0 + kk < K) {
#pragma unroll
                    for (int y = 0; y < YC; ++y) {
                        const T yv = ylds[y * K + k0 + kk];
#pragma unroll
                        for (int i = 0; i < VF; ++i) acc[i][y] = fma(xv[kk][i], yv, acc[i][y]);
                    }
                }
            }
        }
        // outputs (f, y) of this thread: one run of VF * YC entries at f0 * YC
        constexpr int RUN = VF * YC;
        float m = 0.0f;
        if constexpr (RUN * sizeof(T) >= 64) {
            // Runs of >= 64 bytes: direct 16-byte stores of one instruction would land a
            // run apart.  The wave's runs (consecutive in C) go through its 4 KB of LDS, L
            // lanes' runs per round, and each round leaves as 4 stores of 16 bytes per lane
            // over 1 KB of consecutive bytes.
            constexpr int E = 4096 / sizeof(T);  // entries per round
            constexpr int L = E / RUN;           // lanes per round
            static_assert(L >= 1 && L <= 64 && 64 % L == 0, "a round stages whole runs of one wave");
            const int lane = tid & 63;
            T* st = reinterpret_cast<T*>(smem + ((8 * K + 15) & ~15) + (size_t)YC * K * sizeof(T)) + (tid >> 6) * E;
            auto Cw = C + batch * cb + ((((ft << tb) | ((int64_t)(tid & ~63) << LVF))) * YC);
#pragma unroll
            for (int q = 0; q < 64 / L; ++q) {
                if (lane / L == q) {
#pragma unroll
                    for (int j = 0; j < RUN; j += VF) {
                        V o;
#pragma unroll
                        for (int c = 0; c < VF; ++c) {
                            const T v = acc[(j + c) / YC][(j + c) % YC];
                            o[c] = v;
                            if (SCALED) m = fmaxf(m, (float)absval(v));
                        }
                        *reinterpret_cast<V*>(st + (lane % L) * RUN + j) = o;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
                for (int r = 0; r < E / (64 * VF); ++r) {
                    const int e = (r * 64 + lane) * VF;
                    __builtin_nontemporal_store(*reinterpret_cast<const V*>(st + e),
                                                reinterpret_cast<TO_GLOBAL V*>(Cw + q * E + e));
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        } else {
            auto Cp = C + batch * cb + ((((ft << tb) | ((int64_t)tid << LVF))) * YC);
#pragma unroll
            for (int j = 0; j < RUN; j += VF) {
                V o;
#pragma unroll
                for (int c = 0; c < VF; ++c) {
                    const T v = acc[(j + c) / YC][(j + c) % YC];
                    o[c] = v;
                    if (SCALED) m = fmaxf(m, (float)absval(v));
                }
                __builtin_nontemporal_store(o, reinterpret_cast<TO_GLOBAL V*>(Cp + j));
            }
        }
        if (SCALED) {
            mrun = fmaxf(mrun, m);
            if (tid == 0 && ft == 0) note_exp(d->cs, batch, ex + ey);
            if (t + 1 == t1 || ft + 1 == nft) {
                const float w = wave_max(mrun);
                if ((tid & 63) == 0) note_max(d->cs, batch, w);
                mrun = 0.0f;
            }
        }
        if (++ft == nft) {
            ft = 0;
            xhi = 0;
            ++batch;
        } else {
            xhi = ((xhi | ~hmask) + 1) & hmask;
        }
    }
}

template <typename T, bool SCALED, int YC>
__global__ __launch_bounds__(kBlock) void k_join_stream_vec(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_stream_vec_body<T, SCALED, YC>(d, blockIdx.x - d->block0, smem);
}

// ---------------------------------------------------------------------------------
// streaming join, row form: outputs C[f][y] with the 2^fy values of y contiguous
// (shy == 0, shf == fy, fy >= 5) and few shared bits (s <= 4) -- an outer product of
// long rows, bound by the output writes.  Each thread owns VY = 16 bytes of consecutive y
// of a row and keeps their Y[y][k] (VY * 2^S values) in registers for its batch item;
// the threads of one row read the same X[f][k] (one load per wave), form VY outputs and
// write them as one 16-byte store, so a wave writes whole rows.  A block walks a
// contiguous range of tiles (64 rows, or 8 passes of rows if fewer) of one batch item.

template <typename T, bool SCALED>
__device__ __forceinline__ void join_rows_body(const to_join_desc* __restrict__ d, const int local) {
    constexpr int KMAX = 16;  // s <= 4; one kernel for every s, so joins of one batch step
                              // with different s share a launch
    constexpr int VY = 16 / sizeof(T);

    typedef typename Vec16<T>::type V;
    const int tid = threadIdx.x;
    const int fx = d->fx, fy = d->fy;
    const int tpr = (1 << fy) / VY;          // threads per row (the host keeps it <= kBlock)
    const int rb = kBlock / tpr;             // rows per pass
    const int j = tid % tpr, r = tid / tpr;  // y vector and row of this thread
    const int S = d->s, K = 1 << S;
    const TO_GLOBAL T* __restrict__ X = gptr<const T>(d->x);
    const TO_GLOBAL T* __restrict__ Y = gptr<const T>(d->y);
    TO_GLOBAL T* __restrict__ C = gptr<T>(d->c);
    const LaneBits XF = lane_bits(d->xf_pos, fx), XS = lane_bits(d->xs_pos, S);
    const LaneBits YF = lane_bits(d->yf_pos, fy), YS = lane_bits(d->ys_pos, S);
    int64_t kx[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) kx[k] = k < K ? uniform64(pdep_l(k, XS, 0, S)) : 0;
    // a tile is 64 rows, or 8 passes of rb rows when that is fewer: a block's passes are
    // sequential (each waits for its X loads), so wide rows get short tiles, more blocks
    const int lr = min(min(fx, 6), 11 + (sizeof(T) == 4 ? 2 : 1) - fy);  // log2 rows per tile
    const int64_t ROWS = 1ll << lr;
    const int tb = fx - lr;                  // tile bits of f
    const int64_t nft = 1ll << tb;
    const int64_t ntiles = (int64_t)d->nbatch << tb;
    const int64_t t0 = ntiles * local / d->nblocks, t1 = ntiles * (local + 1) / d->nblocks;
    const int64_t fmask = pdep_l(~0ull, XF, 0, fx);
    const int64_t xstep = pdep_l((uint64_t)rb, XF, 0, fx);  // rb rows on, in X's bit layout
    T yr[KMAX][VY];
    int64_t staged = -1;
    float sxy = 1.0f;
    int64_t exy = 0;
    float mrun = 0.0f;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t batch = t >> tb, ft = t & (nft - 1);
        if (batch != staged) {
            if (SCALED) {
                if (staged >= 0) {
                    const float w = wave_max(mrun);
                    if ((tid & 63) == 0) note_max(d->cs, staged, w);
                    mrun = 0.0f;
                }
                const OpScale a = op_scale(d->xs, d->xsb, batch), c = op_scale(d->ys, d->ysb, batch);
                sxy = a.f * c.f;
                exy = a.e + c.e;
            }
            const TO_GLOBAL T* Yb = Y + batch * d->yb;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < K) {
                    const int64_t ky = pdep_l(k, YS, 0, S);
#pragma unroll
                    for (int e = 0; e < VY; ++e) yr[k][e] = Yb[pdep_l(j * VY + e, YF, 0, fy) + ky];
                }
            }
            staged = batch;
        }
        if (SCALED && ft == 0 && tid == 0) note_exp(d->cs, batch, exy);
        const TO_GLOBAL T* Xb = X + batch * d->xb;
        TO_GLOBAL T* Cb = C + batch * d->cb;
        const int64_t f0 = ft * ROWS + r;
        int64_t xo = pdep_l((uint64_t)f0, XF, 0, fx);
        for (int64_t f = f0; f < ft * ROWS + ROWS; f += rb) {
            T xv[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) xv[k] = k < K ? Xb[xo + kx[k]] : T(0);
            V o;
#pragma unroll
            for (int e = 0; e < VY; ++e) {
                T a = T(0);
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (k < K) a = fma(xv[k], yr[k][e], a);
                o[e] = SCALED ? a * T(sxy) : a;
                if (SCALED) mrun = fmaxf(mrun, (float)absval(o[e]));
            }
            *reinterpret_cast<TO_GLOBAL V*>(Cb + (f << fy) + j * VY) = o;
            xo = ((xo | ~fmask) + xstep) & fmask;
        }
    }
    if (SCALED && staged >= 0) {
        const float w = wave_max(mrun);
        if ((tid & 63) == 0) note_max(d->cs, staged, w);
    }
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_rows(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    join_rows_body<T, SCALED>(d, blockIdx.x - d->block0);
}

// Streamed joins of mixed variants in one launch (the many-instance batch groups the joins
// of one step this way): each block runs the form its descriptor's kvar names.
template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_stream_any(const to_join_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const to_join_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int local = blockIdx.x - d->block0;
    switch (d->kvar) {
#define TO_ANY_CASE(YC)                                                  \
    case YC:                                                             \
        join_stream_body<T, SCALED, YC>(d, local, smem);                 \
        break;                                                           \
    case TO_STREAM_VEC | YC:                                             \
        join_stream_vec_body<T, SCALED, YC>(d, local, smem);             \
        break;
        TO_ANY_CASE(1)
        TO_ANY_CASE(2)
        TO_ANY_CASE(4)
        TO_ANY_CASE(8)
        TO_ANY_CASE(16)
#undef TO_ANY_CASE
        case TO_STREAM_ROWS:
            join_rows_body<T, SCALED>(d, local);
            break;
        default:
            break;  // rejected on the host (schedule.stream_variant)
    }
}

// Streamed joins small enough for one wave each (TO_STREAM_WAVE; the many-instance batch
// has tens of thousands of them per step): descriptor i is run by wave i % 4 of block i / 4,
// with no LDS and no barrier, so a block costs four joins one descriptor read each instead
// of one join a table set-up.  The bit deposits are computed once per join, one per lane:
// lane t holds the X / Y offset that bit t of the output index contributes, and lane k the
// offsets of shared index k; the loops read them back with v_readlane (uniform index).  Lanes
// walk the outputs of each batch item with output bits 0-5 on the lane id; bits 0.. of the
// output index are y when shy == 0 (f otherwise).  Requires s <= 6 and ranks of X, Y and
// the output below 32.
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_join_wave(const to_join_desc* __restrict__ descs, int ndesc) {
    const int idx = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    if (idx >= ndesc) return;
    const to_join_desc* d = descs + idx;
    const int lane = threadIdx.x & 63;
    const int fx = d->fx, fy = d->fy, s = d->s;
    const int K = 1 << s;
    const int nbits = fx + fy;
    const int nout = 1 << nbits;
    const bool yinner = d->shy == 0;
    const int shf = d->shf, shy = d->shy;
    // lane t < nbits: what output-index bit t adds to the X offset, the Y offset, and to the
    // output's f and y (as a bit of f << shf | y << shy, which is the output offset)
    uint32_t cx = 0, cy = 0, cc = 0;
    if (lane < nbits) {
        const bool ybit = yinner ? lane < fy : lane >= fx;
        const int b = ybit ? (yinner ? lane : lane - fx) : (yinner ? lane - fy : lane);
        if (ybit) {
            cy = 1u << d->yf_pos[b];
            cc = 1u << (b + shy);
        } else {
            cx = 1u << d->xf_pos[b];
            cc = 1u << (b + shf);
        }
    }
    // lane k < K: the offsets of shared index k in X and in Y
    uint32_t kxl = 0, kyl = 0;
    if (lane < K) {
        kxl = (uint32_t)pdep(lane, d->xs_pos, s);
        kyl = (uint32_t)pdep(lane, d->ys_pos, s);
    }
    // Fewer than 64 outputs: 2^lsub lanes share an output, each summing a contiguous share
    // of K (the share's high k bits are the lane's q, its low bits walk the loop), and the
    // shares are added across those lanes.  Lane = (output within the pass << lsub) | q.
    const int lsub = nbits < 6 ? min(s, 6 - nbits) : 0;
    const int q = lane & ((1 << lsub) - 1), ol = lane >> lsub;
    const int KL = K >> lsub;                 // terms per lane
    const int obits = 6 - lsub;               // output bits per pass
    const uint32_t xq = (uint32_t)__shfl((int)kxl, q << (s - lsub), 64);
    const uint32_t yq = (uint32_t)__shfl((int)kyl, q << (s - lsub), 64);
    // the offsets of this lane's own output bits (the low obits of the output index)
    uint32_t xlane = 0, ylane = 0, clane = 0;
    const int nlow = nbits < obits ? nbits : obits;
    for (int t = 0; t < nlow; ++t) {
        const uint32_t a = lane_value(cx, t), b = lane_value(cy, t), c = lane_value(cc, t);
        if ((ol >> t) & 1) {
            xlane |= a;
            ylane |= b;
            clane |= c;
        }
    }
    const TO_GLOBAL T* __restrict__ X = gptr<const T>(d->x);
    const TO_GLOBAL T* __restrict__ Y = gptr<const T>(d->y);
    TO_GLOBAL T* __restrict__ C = gptr<T>(d->c);
    for (int64_t b = 0; b < d->nbatch; ++b) {
        float sx = 1.0f, sy = 1.0f;
        int64_t e = 0;
        if (SCALED) {
            const OpScale a = op_scale(d->xs, d->xsb, b);
            const OpScale c = op_scale(d->ys, d->ysb, b);
            sx = a.f; sy = c.f; e = a.e + c.e;
        }
        const auto Xb = X + b * d->xb;
        const auto Yb = Y + b * d->yb;
        const auto Cb = C + b * d->cb;
        float m = 0.0f;
        for (int j = 0; (j << obits) < nout; ++j) {
            // output bits obits.. are those of j, the same in every lane
            uint32_t xj = 0, yj = 0, cj = 0;
            for (int t = obits; t < nbits; ++t) {
                if ((j >> (t - obits)) & 1) {
                    xj |= lane_value(cx, t);
                    yj |= lane_value(cy, t);
                    cj |= lane_value(cc, t);
                }
            }
            const bool valid = (j << obits) + ol < nout;
            Acc<T> acc = 0.0;
            if (valid) {
                const auto Xp = Xb + (xlane | xj | xq);
                const auto Yp = Yb + (ylane | yj | yq);
#pragma unroll 4
                for (int k = 0; k < KL; ++k)
                    acc += (Acc<T>)(Xp[lane_value(kxl, k)] * T(sx)) * (Acc<T>)(Yp[lane_value(kyl, k)] * T(sy));
            }
            // every lane takes part in the sums across an output's lanes
            for (int off = 1; off < (1 << lsub); off <<= 1) acc += __shfl_xor(acc, off, 64);
            if (valid && q == 0) {
                const T v = (T)acc;
                Cb[clane | cj] = v;
                if (SCALED) m = fmaxf(m, (float)absval(v));
            }
        }
        if (SCALED) {
            m = wave_max(m);
            if (lane == 0) {
                note_exp(d->cs, b, e);
                note_max(d->cs, b, m);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// bit-permutation of all-dimension-2 tensors, LDS-tiled: a tile spans the union U of
// the m lowest input bits and the input bits that land on the m lowest output bits
// (up to 2^12 entries), so the tile is read with the input's low bits contiguous and
// written with the output's low bits contiguous.  Each thread moves up to kPermEpt
// entries of a tile; the next tile's loads are issued before the current tile is stored.

constexpr int kPermEpt = 16;  // 2^12 tile entries / 256 threads

// LDS address of tile entry u: the tile bits that the 6 lowest output bits come from
// (the lanes of a storing wave vary exactly those) are folded by xor into the low 6
// address bits they do not already cover, so loading lanes (consecutive u) and storing
// lanes both hit distinct banks.  The map is linear over xor and a bijection of the tile.
struct PermSwizzle {
    int n;
    int src[6], dst[6];
};

__device__ inline PermSwizzle perm_swizzle(const to_perm_desc* d) {
    const int nb = d->nu < 6 ? d->nu : 6;
    int used = 0;
    for (int j = 0; j < nb; ++j)
        if (d->v2u[j] < nb) used |= 1 << d->v2u[j];
    PermSwizzle z;
    z.n = 0;
    int pos = 0;
    for (int j = 0; j < nb; ++j) {
        const int b = d->v2u[j];
        if (b >= nb) {
            while ((used >> pos) & 1) ++pos;
            z.src[z.n] = b;
            z.dst[z.n] = pos;
            used |= 1 << pos;
            ++z.n;
        }
    }
    return z;
}

__device__ inline int perm_swz(int u, const PermSwizzle& z) {
    int a = u;
    for (int i = 0; i < z.n; ++i) a ^= ((u >> z.src[i]) & 1) << z.dst[i];
    return a;
}

template <typename T>
__device__ __forceinline__ void permute_tiles(const to_perm_desc* __restrict__ d, int local, int nblocks,
                                              unsigned char* smem) {
    const int nu = d->nu, U = 1 << nu;
    T* tile = reinterpret_cast<T*>(smem);
    int64_t* lo_in = reinterpret_cast<int64_t*>(smem + ((sizeof(T) << nu) + 15) / 16 * 16);
    int64_t* lo_out = lo_in + 256;
    const int nrest = d->nrest;
    const int nlo = nrest < 8 ? nrest : 8;
    for (int r = threadIdx.x; r < (1 << nlo); r += kBlock) {
        lo_in[r] = pdep(r, d->rest_in, nlo);
        lo_out[r] = pdep(r, d->rest_out, nlo);
    }
    const PermSwizzle z = perm_swizzle(d);
    // entry e = tid + 256 i of a tile: its input offset, output offset and LDS addresses
    // are linear in the bits of e (sums / xors over disjoint bits), so each splits into a
    // per-thread part (tid) and a part of i shared by the block
    auto umap = [&](int e) {
        int u = 0;
        for (int j = 0; j < nu; ++j) u |= ((e >> j) & 1) << d->v2u[j];
        return u;
    };
    const int tid = threadIdx.x;
    const int in_lo = (int)pdep(tid, d->u_in, nu), out_lo = (int)pdep(tid, d->v_out, nu);
    const int wa_lo = perm_swz(tid, z), ra_lo = perm_swz(umap(tid), z);
    // images of the 4 bits of i (entry bits 8..11); the part of entry i is their sum/xor
    int bin[4], bout[4], bwa[4], bra[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = kBlock << k;
        bin[k] = (int)pdep(e, d->u_in, nu);
        bout[k] = (int)pdep(e, d->v_out, nu);
        bwa[k] = perm_swz(e, z);
        bra[k] = perm_swz(umap(e), z);
    }
    auto hi_sum = [](const int* b, int i) {
        return ((i & 1) ? b[0] : 0) + ((i & 2) ? b[1] : 0) + ((i & 4) ? b[2] : 0) + ((i & 8) ? b[3] : 0);
    };
    auto hi_xor = [](const int* b, int i) {
        return ((i & 1) ? b[0] : 0) ^ ((i & 2) ? b[1] : 0) ^ ((i & 4) ? b[2] : 0) ^ ((i & 8) ? b[3] : 0);
    };
    __syncthreads();
    const int64_t ntiles = (int64_t)d->nbatch << nrest;
    const int64_t t0 = ntiles * local / nblocks, t1 = ntiles * (local + 1) / nblocks;
    const TO_GLOBAL T* in = gptr<const T>(d->in);
    TO_GLOBAL T* out = gptr<T>(d->out);
    const int64_t rmask = (1ll << nrest) - 1;
    const int lmask = (1 << nlo) - 1;
    // base pointers of tile t (the same for the whole block, kept in scalar registers)
    auto bases = [&](int64_t t, const TO_GLOBAL T*& ip, TO_GLOBAL T*& op) {
        const int64_t b = t >> nrest;
        const int64_t r = t & rmask;
        const int64_t rh = r >> nlo;
        int64_t ib = b * d->inb + pdep(rh, d->rest_in + nlo, nrest - nlo) + lo_in[r & lmask];
        int64_t ob = b * d->outb + pdep(rh, d->rest_out + nlo, nrest - nlo) + lo_out[r & lmask];
        ib = uniform64(ib);
        ob = uniform64(ob);
        ip = in + ib;
        op = out + ob;
    };
    auto run = [&](auto full, auto off32) {
        constexpr bool FULL = decltype(full)::value;
        // in-tile offsets below 2^32 bytes address as 32-bit offsets from the scalar base
        using Off = typename std::conditional<decltype(off32)::value, uint32_t, uint64_t>::type;
        auto live = [&](int i) { return FULL || tid + kBlock * i < U; };
        auto at = [](const TO_GLOBAL T* p, int e) {
            return *reinterpret_cast<const TO_GLOBAL T*>(reinterpret_cast<const TO_GLOBAL char*>(p) +
                                                        (Off)e * (Off)sizeof(T));
        };
        auto to = [](TO_GLOBAL T* p, int e) -> TO_GLOBAL T& {
            return *reinterpret_cast<TO_GLOBAL T*>(reinterpret_cast<TO_GLOBAL char*>(p) + (Off)e * (Off)sizeof(T));
        };
        T val[kPermEpt];
        if (t0 < t1) {
            const TO_GLOBAL T* ip;
            TO_GLOBAL T* op;
            bases(t0, ip, op);
#pragma unroll
            for (int i = 0; i < kPermEpt; ++i)
                if (live(i)) val[i] = at(ip, in_lo + hi_sum(bin, i));
        }
        for (int64_t t = t0; t < t1; ++t) {
            const TO_GLOBAL T* ip;
            TO_GLOBAL T* op;
            bases(t, ip, op);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kPermEpt; ++i)
                if (live(i)) tile[wa_lo ^ hi_xor(bwa, i)] = val[i];
            __syncthreads();
            if (t + 1 < t1) {
                const TO_GLOBAL T* ip2;
                TO_GLOBAL T* op2;
                bases(t + 1, ip2, op2);
#pragma unroll
                for (int i = 0; i < kPermEpt; ++i)
                    if (live(i)) val[i] = at(ip2, in_lo + hi_sum(bin, i));
            }
#pragma unroll
            for (int i = 0; i < kPermEpt; ++i)
                if (live(i)) to(op, out_lo + hi_sum(bout, i)) = tile[ra_lo ^ hi_xor(bra, i)];
        }
    };
    // full tiles (every entry of every thread live) take the path without guards
    const bool off32 = ((uint64_t)sizeof(T) << d->rank) <= (1ull << 32);
    if (U == kBlock * kPermEpt) {
        if (off32) run(std::true_type{}, std::true_type{});
        else run(std::true_type{}, std::false_type{});
    } else {
        run(std::false_type{}, std::false_type{});
    }
}

// Vectorised tiles: 16-byte chunks (the lowest input bits) are loaded and the lowest
// output bits are stored as 16-byte chunks.  The tile has exactly NU bits: kPermVecBits<T>
// (32 KB, 8 chunks per thread) or one more (64 KB, 16 chunks per thread, for tensors of at
// least that many bits: the runs read and written are twice as long, 0.67-0.75 of HBM peak
// on rank-26 float tensors against 0.49-0.53 with 32 KB tiles).  The LDS swizzle only
// permutes whole chunks, folding tile bits that the storing lanes vary into the bank bits
// above the chunk.  Outputs leave with streaming (nontemporal) stores.

template <typename T>
constexpr int kPermVecBits = sizeof(T) == 4 ? 13 : 12;

template <typename T, int NU>
__device__ __forceinline__ void permute_tiles_vec(const to_perm_desc* __restrict__ d, int local, int nblocks,
                                                  unsigned char* smem) {
    constexpr int VL = 16 / sizeof(T);            // entries per chunk
    constexpr int CB = sizeof(T) == 4 ? 2 : 1;    // log2(VL)
    constexpr int NB = sizeof(T) == 4 ? 6 : 5;    // entry-address bits that select the bank
    constexpr int NCH = (1 << NU) / VL / kBlock;  // chunks per thread per tile (8 or 16)
    typedef typename Vec16<T>::type V;
    T* tile = reinterpret_cast<T*>(smem);
    int64_t* lo_in = reinterpret_cast<int64_t*>(smem + (sizeof(T) << NU));
    int64_t* lo_out = lo_in + 256;
    const int nrest = d->nrest;
    const int nlo = nrest < 8 ? nrest : 8;
    for (int r = threadIdx.x; r < (1 << nlo); r += kBlock) {
        lo_in[r] = pdep(r, d->rest_in, nlo);
        lo_out[r] = pdep(r, d->rest_out, nlo);
    }
    // swizzle: storing lanes vary v bits CB..CB+5, i.e. tile bits v2u[CB..CB+5]; those
    // above the bank bits are xor-folded into free bank bits at or above CB
    int zsrc[6], zdst[6], zn = 0;
    {
        int used = 0;
        for (int j = CB; j < CB + 6; ++j)
            if (d->v2u[j] < NB) used |= 1 << d->v2u[j];
        int pos = CB;
        for (int j = CB; j < CB + 6; ++j) {
            const int b = d->v2u[j];
            if (b < NB) continue;
            while (pos < NB && ((used >> pos) & 1)) ++pos;
            if (pos >= NB) break;
            zsrc[zn] = b;
            zdst[zn] = pos;
            used |= 1 << pos;
            ++zn;
        }
    }
    auto swz = [&](int u) {
        int a = u;
        for (int i = 0; i < zn; ++i) a ^= ((u >> zsrc[i]) & 1) << zdst[i];
        return a;
    };
    auto umap = [&](int e) {
        int u = 0;
        for (int j = 0; j < NU; ++j) u |= ((e >> j) & 1) << d->v2u[j];
        return u;
    };
    // chunk q = tid + 256 i covers entries q*VL + c; every address below is linear in the
    // bits of (tid, i, c), so it splits into a per-thread part and parts shared by the block
    const int tid = threadIdx.x;
    const int in_lo = (int)pdep(tid * VL, d->u_in, NU), out_lo = (int)pdep(tid * VL, d->v_out, NU);
    const int wa_lo = swz(tid * VL), ra_lo = swz(umap(tid * VL));
    constexpr int NHB = NCH == 16 ? 4 : 3;      // bits of the chunk index i
    int bin[NHB], bout[NHB], bwa[NHB], bra[NHB], rc[VL];
#pragma unroll
    for (int k = 0; k < NHB; ++k) {
        const int e = (kBlock * VL) << k;
        bin[k] = (int)pdep(e, d->u_in, NU);
        bout[k] = (int)pdep(e, d->v_out, NU);
        bwa[k] = swz(e);
        bra[k] = swz(umap(e));
    }
#pragma unroll
    for (int c = 0; c < VL; ++c) rc[c] = swz(umap(c));
    auto hsum = [](const int* b, int i) {
        return ((i & 1) ? b[0] : 0) + ((i & 2) ? b[1] : 0) + ((i & 4) ? b[2] : 0) + ((NHB > 3 && (i & 8)) ? b[NHB - 1] : 0);
    };
    auto hxor = [](const int* b, int i) {
        return ((i & 1) ? b[0] : 0) ^ ((i & 2) ? b[1] : 0) ^ ((i & 4) ? b[2] : 0) ^ ((NHB > 3 && (i & 8)) ? b[NHB - 1] : 0);
    };
    __syncthreads();
    const int64_t ntiles = (int64_t)d->nbatch << nrest;
    const int64_t t0 = ntiles * local / nblocks, t1 = ntiles * (local + 1) / nblocks;
    const TO_GLOBAL T* in = gptr<const T>(d->in);
    TO_GLOBAL T* out = gptr<T>(d->out);
    const int64_t rmask = (1ll << nrest) - 1;
    const int lmask = (1 << nlo) - 1;
    auto bases = [&](int64_t t, const TO_GLOBAL T*& ip, TO_GLOBAL T*& op) {
        const int64_t b = t >> nrest;
        const int64_t r = t & rmask;
        const int64_t rh = r >> nlo;
        int64_t ib = b * d->inb + pdep(rh, d->rest_in + nlo, nrest - nlo) + lo_in[r & lmask];
        int64_t ob = b * d->outb + pdep(rh, d->rest_out + nlo, nrest - nlo) + lo_out[r & lmask];
        ip = in + uniform64(ib);
        op = out + uniform64(ob);
    };
    V val[NCH];
    if (t0 < t1) {
        const TO_GLOBAL T* ip;
        TO_GLOBAL T* op;
        bases(t0, ip, op);
#pragma unroll
        for (int i = 0; i < NCH; ++i) val[i] = *reinterpret_cast<const TO_GLOBAL V*>(ip + (in_lo + hsum(bin, i)));
    }
    for (int64_t t = t0; t < t1; ++t) {
        const TO_GLOBAL T* ip;
        TO_GLOBAL T* op;
        bases(t, ip, op);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NCH; ++i) *reinterpret_cast<V*>(tile + (wa_lo ^ hxor(bwa, i))) = val[i];
        __syncthreads();
        if (t + 1 < t1) {
            const TO_GLOBAL T* ip2;
            TO_GLOBAL T* op2;
            bases(t + 1, ip2, op2);
#pragma unroll
            for (int i = 0; i < NCH; ++i)
                val[i] = *reinterpret_cast<const TO_GLOBAL V*>(ip2 + (in_lo + hsum(bin, i)));
        }
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int ra = ra_lo ^ hxor(bra, i);
            V o;
#pragma unroll
            for (int c = 0; c < VL; ++c) o[c] = tile[ra ^ rc[c]];
            // streaming stores: the permuted tensor is read once, later, by another kernel
            __builtin_nontemporal_store(o, reinterpret_cast<TO_GLOBAL V*>(op + (out_lo + hsum(bout, i))));
        }
    }
}

template <typename T, int NU>
__global__ __launch_bounds__(kBlock) void k_permute_vec_desc(const to_perm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const to_perm_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    permute_tiles_vec<T, NU>(d, blockIdx.x - d->block0, d->nblocks, smem);
}

template <typename T, int NU>
__global__ __launch_bounds__(kBlock) void k_permute_vec(const to_perm_desc desc) {
    extern __shared__ __align__(16) unsigned char smem[];
    permute_tiles_vec<T, NU>(&desc, blockIdx.x, desc.nblocks, smem);
}

// descriptor-driven bit permutation, used to bring the two operands of a dense join into
// row-major [free][shared] form for the MFMA product
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute_desc(const to_perm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const to_perm_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    permute_tiles<T>(d, blockIdx.x - d->block0, d->nblocks, smem);
}

// the same permutation with its descriptor passed by value (to_permute)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute(const to_perm_desc desc) {
    extern __shared__ __align__(16) unsigned char smem[];
    permute_tiles<T>(&desc, blockIdx.x, desc.nblocks, smem);
}

// ---------------------------------------------------------------------------------
// MFMA product of a dense join: C[m][n] = sum_k A[m][k] * B[n][k] on 64x64 output tiles.
// 4 waves per block, each owning a 32x32 quarter as 2x2 MFMA 16x16 tiles; K advances 16 at
// a time through LDS (k-major, so the 16 lanes reading one MFMA operand row hit 16
// consecutive LDS words).  Operand layout of the 16x16x4 instructions: A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k; the result D[i][j] is in lane j + 16(i/4), element
// i%4 for f32, and in lane j + 16(i%4), element i/4 for f64 (Mfma<T>::row).

// Workgroups are dealt to the 8 XCDs round-robin, and each XCD has its own L2.  Tile
// index t (block order) is mapped so that the tiles one XCD runs are consecutive in
// (batch, tm, tn) order and share their A panel (same tm) in that XCD's L2: the blocks
// t = 8q + x run tile x * (ntiles / 8) + q.  A bijection of [0, ntiles) whenever ntiles
// is a multiple of 8, the identity otherwise.
__device__ __forceinline__ int64_t xcd_tile(int64_t t, int64_t ntiles) {
    if (ntiles & 7) return t;
    return (t & 7) * (ntiles >> 3) + (t >> 3);
}

// Tile (tm, tn) of the rem-th tile of one batch item, in groups of 8 tile rows: the
// tiles of a group are taken column by column (8 row tiles, then the next column), so the
// blocks running at once on one XCD cover about an 8x8 patch and share 8 A and 8 B
// panels in its L2, instead of one A panel and every B panel of a row.  A bijection of
// [0, tiles_m * tiles_n) (both powers of two).
__device__ __forceinline__ void grouped_tile(int64_t rem, int64_t tiles_m, int64_t tiles_n, int64_t& tm,
                                             int64_t& tn) {
    const int64_t g = tiles_m < 8 ? tiles_m : 8;
    const int64_t span = g * tiles_n;
    const int64_t group = rem / span, r = rem - group * span;
    tm = group * g + (r & (g - 1));
    tn = r / g;
}

// A block walks its tiles in order and keeps one running max |output| per wave for the
// batch item of the current tile; it is published (one atomic per wave) when the batch
// item changes and after the last tile, rather than once per tile.  `next` is the batch
// item of the coming tile (-1 after the last); wave-uniform, as the cross-lane max needs.
__device__ __forceinline__ void publish_max_on_change(uint64_t recs, int64_t& cur, int64_t next, float& mx,
                                                      int lane) {
    if (next == cur) return;
    if (cur >= 0) {
        const float w = wave_max(mx);
        if (lane == 0) note_max(recs, cur, w);
    }
    cur = next;
    mx = 0.0f;
}

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V op(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of result element v held by `lane`
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) * 4 + v; }
};

// With WIDE, float products are summed by the MFMA over 64 (64x64 tiles) or 128 (128-row
// tiles) terms and those partial sums are accumulated in double, so a long shared range (K up to 2^15 here)
// does not grow the float rounding error with sqrt(K); it costs registers (occupancy), so
// the host asks for it only for long shared ranges.
template <typename T, bool WIDE>
constexpr bool kWideAccum = WIDE && sizeof(T) == 4;

template <typename T, int N, int M>
__device__ __forceinline__ void flush_wide(typename Mfma<T>::V (&acc)[N][M], double (&wide)[N][M][4]) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                wide[i][j][v] += (double)acc[i][j][v];
                acc[i][j][v] = T(0);
            }
}

template <typename T, bool SCALED, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_gemm(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    // K slabs of 32: each thread fetches 8 consecutive k of one A row and one B row as
    // 16-byte loads, one slab ahead (the next slab's loads are in flight while this slab's
    // MFMAs run); K = 16 fills the slab's upper half with zeros
    constexpr int BM = 64, BN = 64, BK = 32, PAD = 4, KT = 8;
    constexpr int VE = 16 / sizeof(T);
    typedef typename Vec16<T>::type V16;
    __shared__ T As[BK][BM + PAD];
    __shared__ T Bs[BK][BN + PAD];
    typedef typename Mfma<T>::V V;
    const to_gemm_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t K = 1ll << d->klog;
    const int64_t tiles_m = 1ll << (d->mlog - 6), tiles_n = 1ll << (d->nlog - 6);
    const int64_t per_batch = tiles_m * tiles_n;
    // split-K (lsplit > 0): tile t's K range is cut into 2^lsplit parts, one block each,
    // whose double partial sums go to c (as double[outputs][2^lsplit]) for a
    // TO_OP_JOIN_FINISH launch to add up, scale-record updates included
    const int lsplit = d->lsplit;
    const int64_t KS = K >> lsplit;  // this block's share of K
    const int64_t ntiles = ((int64_t)d->nbatch * per_batch) << lsplit;
    const int lrow = tid >> 2, lk = (tid & 3) * KT;
    const bool kin = lk < KS;  // this thread's 8 k exist (K >= 32, or the lower half of K = 16)
    float mx = 0.0f;        // running max |output| of the batch item mx_batch
    int64_t mx_batch = -1;
    for (int64_t t0 = blockIdx.x - d->block0; t0 < ntiles; t0 += d->nblocks) {
        const int64_t t = xcd_tile(t0, ntiles);
        const int64_t sp = t & ((1ll << lsplit) - 1);
        const int64_t batch = (t >> lsplit) / per_batch;
        const int64_t rem = (t >> lsplit) - batch * per_batch;
        int64_t tm, tn;
        grouped_tile(rem, tiles_m, tiles_n, tm, tn);
        if (SCALED && lsplit == 0) publish_max_on_change(d->cs, mx_batch, batch, mx, lane);
        const TO_GLOBAL T* A = gptr<const T>(d->a) + batch * d->ab + (tm * BM + lrow) * K + sp * KS + lk;
        const TO_GLOBAL T* B = gptr<const T>(d->b) + batch * d->bb + (tn * BN + lrow) * K + sp * KS + lk;
        T sa = T(1), sb = T(1);
        int64_t ea = 0, eb = 0;
        if (SCALED) {
            const OpScale x = op_scale(d->as, d->asb, batch), y = op_scale(d->bs, d->bsb, batch);
            sa = T(x.f); ea = x.e; sb = T(y.f); eb = y.e;
        }
        V acc[2][2];
        double wide[2][2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    acc[i][j][v] = T(0);
                    wide[i][j][v] = 0.0;
                }
        V16 ra[KT / VE], rb[KT / VE];
        auto fetch = [&](int64_t k0) {
#pragma unroll
            for (int j = 0; j < KT / VE; ++j) {
                ra[j] = kin ? *reinterpret_cast<const TO_GLOBAL V16*>(A + k0 + j * VE) : V16{};
                rb[j] = kin ? *reinterpret_cast<const TO_GLOBAL V16*>(B + k0 + j * VE) : V16{};
            }
        };
        fetch(0);
        int64_t step = 0;
        for (int64_t k0 = 0; k0 < KS; k0 += BK, ++step) {
            __syncthreads();  // the previous slab's fragments have been read
#pragma unroll
            for (int j = 0; j < KT / VE; ++j)
#pragma unroll
                for (int e = 0; e < VE; ++e) {
                    As[lk + j * VE + e][lrow] = ra[j][e] * sa;
                    Bs[lk + j * VE + e][lrow] = rb[j][e] * sb;
                }
            __syncthreads();
            if (k0 + BK < KS) fetch(k0 + BK);
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                const int kr = kk * 4 + (lane >> 4);
                T a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
            }
            // partial sums of 64 terms go into the double accumulators
            if (kWideAccum<T, WIDE> && ((step & 1) == 1 || k0 + BK >= KS)) flush_wide<T, 2, 2>(acc, wide);
        }
        if (lsplit > 0) {
            TO_GLOBAL double* P = gptr<double>(d->c);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int64_t n = tn * BN + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int64_t m = tm * BM + wm * 32 + i * 16 + Mfma<T>::row(lane, v);
                        const int64_t o = (batch << (d->mlog + d->nlog)) | (m << d->nlog) | n;
                        P[(o << lsplit) + sp] = kWideAccum<T, WIDE> ? wide[i][j][v] : (double)acc[i][j][v];
                    }
                }
            continue;
        }
        TO_GLOBAL T* C = gptr<T>(d->c) + batch * d->cb;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int64_t n = tn * BN + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t m = tm * BM + wm * 32 + i * 16 + Mfma<T>::row(lane, v);
                    const T out = kWideAccum<T, WIDE> ? T(wide[i][j][v]) : acc[i][j][v];
                    C[m * d->ldcm + n * d->ldcn] = out;
                    if (SCALED) mx = fmaxf(mx, (float)absval(out));
                }
            }
        }
        if (SCALED && tid == 0 && tm == 0 && tn == 0) note_exp(d->cs, batch, ea + eb);
    }
    if (SCALED) publish_max_on_change(d->cs, mx_batch, -1, mx, lane);
}

// 128x128 output tiles: each wave owns a 64x64 quarter as 4x4 MFMA tiles; K advances by
// BK (16 for double, 32 for float) with 16-byte global loads, and the next K slab is
// fetched into registers while the current one is multiplied out of LDS.  With DB the
// LDS slabs are double-buffered: the next slab is stored while the current one is still
// being read, so one barrier per K step suffices (the form for one block per CU, where no
// second block covers a barrier wait).
template <typename T, bool SCALED, bool WIDE, int BN, bool TR, bool DB>
__device__ __forceinline__ void gemm128_body(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    constexpr int BM = 128;
    constexpr int NBUF = DB ? 2 : 1;
    constexpr int NJ = BN / 32;                      // 16-column MFMA tiles per wave
    constexpr int VE = 16 / sizeof(T);              // elements per 16-byte load
    constexpr int BK = sizeof(T) == 8 ? 16 : 32;     // the host requires K >= BK
    constexpr int FLUSH = 128 / BK;                  // K steps per float partial sum (128 terms)
    constexpr int VPR = BK / VE;                     // 16-byte vectors per tile row
    constexpr int NV = BM * VPR / kBlock;            // A vectors per thread
    constexpr int NVB = BN * VPR / kBlock;           // B vectors per thread (4 or 2)
    constexpr int LDK = BK + VE;                     // LDS row pitch: k contiguous, 16-byte pad
    __shared__ __align__(16) T As[NBUF][BM][LDK];
    __shared__ __align__(16) T Bs[NBUF][BN][LDK];
    typedef typename Mfma<T>::V V;
    typedef T L __attribute__((ext_vector_type(VE)));
    const to_gemm_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t K = 1ll << d->klog;
    const int64_t tiles_m = 1ll << (d->mlog - 7), tiles_n = 1ll << (d->nlog - (BN == 128 ? 7 : 6));
    const int64_t per_batch = tiles_m * tiles_n;
    const int64_t ntiles = (int64_t)d->nbatch * per_batch;
    int vrow[NV], vk[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int v = tid + kBlock * i;
        vrow[i] = v / VPR;
        vk[i] = (v % VPR) * VE;
    }
    float mx = 0.0f;        // running max |output| of the batch item mx_batch
    int64_t mx_batch = -1;
    for (int64_t t0 = blockIdx.x - d->block0; t0 < ntiles; t0 += d->nblocks) {
        const int64_t t = xcd_tile(t0, ntiles);
        const int64_t batch = t / per_batch;
        const int64_t rem = t - batch * per_batch;
        int64_t tm, tn;
        grouped_tile(rem, tiles_m, tiles_n, tm, tn);
        if (SCALED) publish_max_on_change(d->cs, mx_batch, batch, mx, lane);
        const TO_GLOBAL T* A = gptr<const T>(d->a) + batch * d->ab + tm * BM * K;
        const TO_GLOBAL T* B = gptr<const T>(d->b) + batch * d->bb + tn * BN * K;
        T sa = T(1), sb = T(1);
        int64_t ea = 0, eb = 0;
        if (SCALED) {
            const OpScale x = op_scale(d->as, d->asb, batch), y = op_scale(d->bs, d->bsb, batch);
            sa = T(x.f); ea = x.e; sb = T(y.f); eb = y.e;
        }
        V acc[4][NJ];
        double wide[4][NJ][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    acc[i][j][v] = T(0);
                    wide[i][j][v] = 0.0;
                }
        int64_t step = 0;
        L ra[NV], rb[NVB];
#pragma unroll
        for (int i = 0; i < NV; ++i) ra[i] = *reinterpret_cast<const TO_GLOBAL L*>(A + vrow[i] * K + vk[i]);
#pragma unroll
        for (int i = 0; i < NVB; ++i) rb[i] = *reinterpret_cast<const TO_GLOBAL L*>(B + vrow[i] * K + vk[i]);
        // rows keep k contiguous, so each 16-byte slab goes to LDS in one store; the
        // operands' power-of-two scales are applied to the outputs instead (exact)
        if (DB) {
            __syncthreads();  // the previous tile's last slab has been read
#pragma unroll
            for (int i = 0; i < NV; ++i) *reinterpret_cast<L*>(&As[0][vrow[i]][vk[i]]) = ra[i];
#pragma unroll
            for (int i = 0; i < NVB; ++i) *reinterpret_cast<L*>(&Bs[0][vrow[i]][vk[i]]) = rb[i];
            __syncthreads();
        }
        for (int64_t k0 = 0; k0 < K; k0 += BK, ++step) {
            const int cur = DB ? (int)(step & 1) : 0;
            const bool more = k0 + BK < K;
            if (!DB) {
                __syncthreads();
#pragma unroll
                for (int i = 0; i < NV; ++i) *reinterpret_cast<L*>(&As[0][vrow[i]][vk[i]]) = ra[i];
#pragma unroll
                for (int i = 0; i < NVB; ++i) *reinterpret_cast<L*>(&Bs[0][vrow[i]][vk[i]]) = rb[i];
                __syncthreads();
            }
            if (more) {
#pragma unroll
                for (int i = 0; i < NV; ++i)
                    ra[i] = *reinterpret_cast<const TO_GLOBAL L*>(A + vrow[i] * K + k0 + BK + vk[i]);
#pragma unroll
                for (int i = 0; i < NVB; ++i)
                    rb[i] = *reinterpret_cast<const TO_GLOBAL L*>(B + vrow[i] * K + k0 + BK + vk[i]);
            }
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                const int kr = kk * 4 + (lane >> 4);
                T a[4], b[NJ];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = As[cur][wm * 64 + i * 16 + (lane & 15)][kr];
#pragma unroll
                for (int j = 0; j < NJ; ++j) b[j] = Bs[cur][wn * (BN / 2) + j * 16 + (lane & 15)][kr];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
                        acc[i][j] = TR ? Mfma<T>::op(b[j], a[i], acc[i][j]) : Mfma<T>::op(a[i], b[j], acc[i][j]);
            }
            if (kWideAccum<T, WIDE> && ((step % FLUSH) == FLUSH - 1 || !more)) flush_wide<T, 4, NJ>(acc, wide);
            if (DB) {
                // the other buffer was last read before the previous barrier
                if (more) {
#pragma unroll
                    for (int i = 0; i < NV; ++i) *reinterpret_cast<L*>(&As[cur ^ 1][vrow[i]][vk[i]]) = ra[i];
#pragma unroll
                    for (int i = 0; i < NVB; ++i) *reinterpret_cast<L*>(&Bs[cur ^ 1][vrow[i]][vk[i]]) = rb[i];
                }
                __syncthreads();
            }
        }

        // Result element v of a lane: without TR, row (m) Mfma::row(lane, v) of column
        // lane & 15; with TR (product computed transposed), column (n) Mfma::row(lane, v)
        // of row lane & 15.  For float the 4 elements are consecutive, so when they are
        // adjacent in C (rows with ldcm == 1, columns with TR and ldcn == 1) one 16-byte
        // store writes them.
        TO_GLOBAL T* C = gptr<T>(d->c) + batch * d->cb;
        const T sab = sa * sb;
        const bool vec_out = sizeof(T) == 4 && (TR ? d->ldcn == 1 : d->ldcm == 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                T out[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    out[v] = (kWideAccum<T, WIDE> ? T(wide[i][j][v]) : acc[i][j][v]) * sab;
                    if (SCALED) mx = fmaxf(mx, (float)absval(out[v]));
                }
                const int64_t mb = tm * BM + wm * 64 + i * 16, nb = tn * BN + wn * (BN / 2) + j * 16;
                if (vec_out) {
                    typedef T V4 __attribute__((ext_vector_type(4)));
                    const int64_t m = TR ? mb + (lane & 15) : mb + Mfma<T>::row(lane, 0);
                    const int64_t n = TR ? nb + Mfma<T>::row(lane, 0) : nb + (lane & 15);
                    *reinterpret_cast<TO_GLOBAL V4*>(C + m * d->ldcm + n * d->ldcn) = V4{out[0], out[1], out[2], out[3]};
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int64_t m = TR ? mb + (lane & 15) : mb + Mfma<T>::row(lane, v);
                        const int64_t n = TR ? nb + Mfma<T>::row(lane, v) : nb + (lane & 15);
                        C[m * d->ldcm + n * d->ldcn] = out[v];
                    }
                }
            }
        }
        if (SCALED && tid == 0 && tm == 0 && tn == 0) note_exp(d->cs, batch, ea + eb);
    }
    if (SCALED) publish_max_on_change(d->cs, mx_batch, -1, mx, lane);
}

// 128x128 tiles with float accumulators, and 128x64 tiles for the wide (double) form,
// both at two waves per SIMD
template <typename T, bool SCALED, bool WIDE, bool TR>
__global__ __launch_bounds__(kBlock, 2) void k_gemm128(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    gemm128_body<T, SCALED, WIDE, WIDE ? 64 : 128, TR, false>(descs, ndesc, bmap);
}

// 128x128 tiles at one block (one wave per SIMD) per CU: the register file of a SIMD
// (accumulation registers included) holds the float and double accumulators of a 64x64
// quarter, so the wide form keeps the 128-column tile; LDS double-buffered
template <typename T, bool SCALED, bool WIDE, bool TR>
__global__ __launch_bounds__(kBlock, 1) void k_gemm128_db(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    gemm128_body<T, SCALED, WIDE, 128, TR, true>(descs, ndesc, bmap);
}

// ---------------------------------------------------------------------------------
// Float products on the bf16 matrix cores (three-way split, six products).
//
// Every float operand x is split into three bf16 parts by round-to-nearest:
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m); both residuals are exact in float and
// x = h + m + l exactly (each step leaves at most 8 significant bits).  Then
//   x * y = hh + (hm + mh) + (hl + mm + lh) + (ml + lm) + ll
// and the six leading products are formed: the three dropped ones are below 2^-23 |x y|
// and have random sign, so a product keeps float accuracy.  The bf16 products are exact
// in the float accumulator of v_mfma_f32_16x16x32_bf16 (16 cycles for 16x16x32), so six
// of them cost 6/16 of one float MFMA of the same depth (16x16x4 in 32 cycles, x8).
// Inputs below 2^-110 of a tensor's largest entry lose bits in the bf16 subnormal range
// (the float path loses them below 2^-126).
//
// 128x128 tiles, 8 waves with a 64x32 part each, K slabs of 32 floats fetched 16 bytes
// at a time two slabs ahead; each float is split once per block when its slab is written
// to LDS, as three bf16 planes [plane][row][k] whose row pitch (40 bf16 = 80 bytes) keeps
// the 16-byte fragment reads conflict-free; LDS double-buffered, one block per CU.  MFMA
// operand layout: lane i + 16g holds k = 8g..8g+7 of row (A) or column (B) i; the result
// layout is that of the other 16x16 instructions.

typedef __bf16 BF8 __attribute__((ext_vector_type(8)));
typedef __bf16 BF2 __attribute__((ext_vector_type(2)));
typedef float F2 __attribute__((ext_vector_type(2)));
typedef uint32_t U4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {  // a in the low half
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(F2{a, b}, BF2));
}

// the three bf16 planes of two floats, each pair packed in one word
__device__ __forceinline__ void split3(float a, float b, uint32_t& h, uint32_t& m, uint32_t& l) {
    h = pack_bf16(a, b);
    a -= __uint_as_float(h << 16);
    b -= __uint_as_float(h & 0xffff0000u);
    m = pack_bf16(a, b);
    a -= __uint_as_float(m << 16);
    b -= __uint_as_float(m & 0xffff0000u);
    l = pack_bf16(a, b);
}

__device__ __forceinline__ Mfma<float>::V mfma_bf16(U4 a, U4 b, Mfma<float>::V c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(BF8, a), __builtin_bit_cast(BF8, b), c, 0,
                                                  0, 0);
}

// Waves per block of the split product: 8, two per SIMD.  With 4 waves (one per SIMD) a
// wave waits on each fragment read with nothing else to issue; 8 waves measured 3-16 %
// faster.
template <bool WIDE>
constexpr int x6_waves() {
    return 8;
}

// Block tiles of BM x 128 outputs: BM = 128 gives each wave a 64x32 part (float or wide
// accumulators); BM = 256 a 64x64 part, which reads 30 % fewer LDS bytes and stages 25 %
// fewer per product and has half the barriers (float accumulators only: the wide form's
// double accumulators of a 64x64 part do not fit two waves per SIMD).
// MFMA/VALU interleaving of one product: QN groups of one MFMA and four VALU, then the
// product's remaining MFMAs
constexpr int x6_groups(int q, int nt, int tj) {
    return 6 * ((q + 1) * nt / 4 - q * nt / 4) < 4 * tj ? 6 * ((q + 1) * nt / 4 - q * nt / 4) : 4 * tj;
}

template <int QN, int TOTAL>
__device__ __forceinline__ void x6_interleave() {
#pragma unroll
    for (int g = 0; g < QN; ++g) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // four VALU
    }
    __builtin_amdgcn_sched_group_barrier(0x008, TOTAL - QN, 0);
}

template <bool SCALED, bool WIDE, bool TR, int BM>
__device__ __forceinline__ void gemm_x6_body(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    constexpr int BN = 128, BK = 32;
    constexpr int NW = x6_waves<WIDE>();             // waves per block
    static_assert(NW == 8 && (BM == 128 || (BM == 256 && !WIDE)), "tile shapes");
    constexpr int WMW = BM / 64;                     // waves along m
    constexpr int WN = NW / WMW;                     // waves along n
    constexpr int TJ = BN / WN / 16;                 // 16-column MFMA tiles per wave (2 or 4)
    constexpr int NVA = BM * (BK / 4) / (64 * NW);   // 16-byte A vectors per thread and slab
    constexpr int NVB = BN * (BK / 4) / (64 * NW);   // B vectors
    constexpr int NT = NVA + NVB;
    constexpr int FLUSH = 4;                         // K steps per float partial sum (128 terms)
    // Chunk-major bf16 planes [k chunk of 8][row][8]: the 16 lanes of a fragment read
    // (rows r..r+15 of one k chunk) take 256 contiguous bytes, which gfx950 serves without
    // bank conflicts (row-major rows padded to 80 bytes gave 2-way conflicts).
    constexpr int PA = (BK / 8) * BM * 16, PB = (BK / 8) * BN * 16;  // bytes per plane
    constexpr int BUF = 3 * (PA + PB);
    typedef Mfma<float>::V V;
    typedef float L __attribute__((ext_vector_type(4)));
    typedef uint32_t U2 __attribute__((ext_vector_type(2)));
    __shared__ __align__(16) unsigned char smem[2 * BUF];  // [buf][A planes][B planes]
    const to_gemm_desc* d = descs + find_desc(descs, ndesc, blockIdx.x, bmap);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int64_t K = 1ll << d->klog;
    const int64_t tiles_m = 1ll << (d->mlog - (BM == 256 ? 8 : 7)), tiles_n = 1ll << (d->nlog - 7);
    const int64_t per_batch = tiles_m * tiles_n;
    const int64_t ntiles = (int64_t)d->nbatch * per_batch;
    // Vector v < NVA is A's vector v, else B's vector v - NVA: 16 bytes (4 k) of row
    // 16 wave + (lane & 15) + 128 (i >> 1), k = 4 (lane >> 4) + 16 (i & 1), so the 32 lanes
    // of one 8-byte store pass write 256 contiguous bytes of one k chunk.
    int vrow[NT], vk[NT], voff[NT];
#pragma unroll
    for (int v = 0; v < NT; ++v) {
        const int i = v < NVA ? v : v - NVA;
        const int rows = v < NVA ? BM : BN;
        vrow[v] = 16 * wave + (lane & 15) + 128 * (i >> 1);
        vk[v] = 4 * (lane >> 4) + 16 * (i & 1);
        voff[v] = (v < NVA ? 0 : 3 * PA) + ((vk[v] >> 3) * rows + vrow[v]) * 16 + (vk[v] & 7) * 2;
    }
    // this lane's fragment (row lane & 15 of a 16-row group, k chunk lane >> 4) in a plane
    const int frag_a = ((lane >> 4) * BM + (lane & 15)) * 16 + (wm * 4) * 256;
    const int frag_b = 3 * PA + ((lane >> 4) * BN + (lane & 15)) * 16 + (wn * TJ) * 256;
    float mx = 0.0f;
    int64_t mx_batch = -1;
    for (int64_t t0 = blockIdx.x - d->block0; t0 < ntiles; t0 += d->nblocks) {
        const int64_t t = xcd_tile(t0, ntiles);
        const int64_t batch = t / per_batch;
        const int64_t rem = t - batch * per_batch;
        int64_t tm, tn;
        grouped_tile(rem, tiles_m, tiles_n, tm, tn);
        if (SCALED) publish_max_on_change(d->cs, mx_batch, batch, mx, lane);
        const TO_GLOBAL float* A = gptr<const float>(d->a) + batch * d->ab + tm * BM * K;
        const TO_GLOBAL float* B = gptr<const float>(d->b) + batch * d->bb + tn * BN * K;
        float sa = 1.0f, sb = 1.0f;
        int64_t ea = 0, eb = 0;
        if (SCALED) {
            const OpScale x = op_scale(d->as, d->asb, batch), y = op_scale(d->bs, d->bsb, batch);
            sa = x.f; ea = x.e; sb = y.f; eb = y.e;
        }
        V acc[4][TJ];
        double wide[WIDE ? 4 : 1][WIDE ? TJ : 1][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    acc[i][j][v] = 0.0f;
                    if (WIDE) wide[WIDE ? i : 0][WIDE ? j : 0][v] = 0.0;
                }
        // Software pipeline over K slabs of 32: while K step s multiplies LDS buffer
        // s & 1, each of the thread's NT fetched vectors of slab s + 1 is split into the
        // other buffer between two products, and its register is refilled at once from
        // slab s + 2.  A vector is thus split at the same point of every step, one step
        // after its fetch, and the splitting (VALU) runs beside the MFMAs.  Fetches past
        // the end re-read the last slab (in bounds; what they stage is never read).  One
        // barrier per step.
        L rv[NT];
        auto fetch_vec = [&](int v, int64_t k) {
            k = k < K - BK ? k : K - BK;
            const TO_GLOBAL float* src = v < NVA ? A : B;
            rv[v] = *reinterpret_cast<const TO_GLOBAL L*>(src + vrow[v] * K + k + vk[v]);
        };
        auto stage_vec = [&](int v, int buf) {
            uint32_t h0, m0, l0, h1, m1, l1;
            split3(rv[v][0], rv[v][1], h0, m0, l0);
            split3(rv[v][2], rv[v][3], h1, m1, l1);
            unsigned char* P = smem + buf * BUF + voff[v];
            const int pl = v < NVA ? PA : PB;
            *reinterpret_cast<U2*>(P) = U2{h0, h1};
            *reinterpret_cast<U2*>(P + pl) = U2{m0, m1};
            *reinterpret_cast<U2*>(P + 2 * pl) = U2{l0, l1};
        };
        // stage the q-th share of the vectors of slab s + 1 into `nbuf` and fetch them from
        // slab s + 2 (at k2)
        auto pipe = [&](int q, int nbuf, int64_t k2) {
#pragma unroll
            for (int v = q * NT / 4; v < (q + 1) * NT / 4; ++v) stage_vec(v, nbuf);
#pragma unroll
            for (int v = q * NT / 4; v < (q + 1) * NT / 4; ++v) fetch_vec(v, k2);
        };
        // The six products of LDS buffer `buf`, the large one last (a float accumulator
        // that takes the small terms first rounds less): (l,h) (m,h) (m,m) (h,m) (h,l) (h,h).
        // Consecutive products share one operand plane, so each product is preceded by the
        // read of at most one new plane, issued while the previous product's MFMAs run: six
        // plane reads for six products (B's high plane stays in registers from the first
        // product to the last).  Scheduling barriers keep the products in order, so at most
        // four planes of fragments are live, not all six.
        auto step = [&](int buf, bool fresh, int64_t k2) {
            const int nbuf = buf ^ 1;
            const unsigned char* S = smem + buf * BUF;
            U4 fa[4], fb[4], fn[4], fh[4];
            auto read_a = [&](U4 (&f)[4], int plane) {
#pragma unroll
                for (int i = 0; i < 4; ++i) f[i] = *reinterpret_cast<const U4*>(S + plane * PA + frag_a + i * 256);
            };
            auto read_b = [&](U4 (&f)[4], int plane) {
#pragma unroll
                for (int j = 0; j < TJ; ++j) f[j] = *reinterpret_cast<const U4*>(S + plane * PB + frag_b + j * 256);
            };
            auto product = [&](const U4 (&a)[4], const U4 (&b)[4]) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j)
                        acc[i][j] = TR ? mfma_bf16(b[j], a[i], acc[i][j]) : mfma_bf16(a[i], b[j], acc[i][j]);
            };
            read_b(fb, 0);   // B h
            read_a(fa, 2);   // A l
            read_a(fn, 1);   // A m (next)
            // Interleave the splitting of the q-th share of the vectors (about 22 VALU
            // each) with the 4 TJ MFMAs of one product: while an MFMA runs (16 cycles),
            // the wave issues the next VALU instructions.
            if (fresh) {     // the first product of a partial sum starts from zero
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j)
                        acc[i][j] = TR ? mfma_bf16(fb[j], fa[i], V{0.0f, 0.0f, 0.0f, 0.0f})
                                       : mfma_bf16(fa[i], fb[j], V{0.0f, 0.0f, 0.0f, 0.0f});
            } else {
                product(fa, fb);                   // (l,h)
            }
            pipe(0, nbuf, k2);
            x6_interleave<x6_groups(0, NT, TJ), 4 * TJ>();
            __builtin_amdgcn_sched_barrier(0);
            read_b(fa, 1);   // B m (next; A l is done)
            product(fn, fb);                       // (m,h)
            pipe(1, nbuf, k2);
            x6_interleave<x6_groups(1, NT, TJ), 4 * TJ>();
            __builtin_amdgcn_sched_barrier(0);
            read_a(fh, 0);   // A h (next; B h stays for the last product)
            product(fn, fa);                       // (m,m)
            pipe(2, nbuf, k2);
            x6_interleave<x6_groups(2, NT, TJ), 4 * TJ>();
            __builtin_amdgcn_sched_barrier(0);
            read_b(fn, 2);   // B l (next; A m is done): the sixth and last plane read
            product(fh, fa);                       // (h,m)
            pipe(3, nbuf, k2);
            x6_interleave<x6_groups(3, NT, TJ), 4 * TJ>();
            __builtin_amdgcn_sched_barrier(0);
            product(fh, fn);                       // (h,l)
            __builtin_amdgcn_sched_barrier(0);
            product(fh, fb);                       // (h,h)
        };
#pragma unroll
        for (int v = 0; v < NT; ++v) fetch_vec(v, 0);
        __syncthreads();  // the previous tile's last slab has been read
#pragma unroll
        for (int v = 0; v < NT; ++v) {
            stage_vec(v, 0);
            fetch_vec(v, BK);
        }
        __syncthreads();
        // unrolled by the partial-sum length when wide (the host requires K >= FLUSH * BK):
        // its first product starts from zero and its last step adds the float partial sums
        // into the double accumulators; by the buffer parity otherwise
        constexpr int UNR = WIDE ? FLUSH : 2;
        for (int64_t k0 = 0; k0 < K; k0 += UNR * BK) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int64_t kk = k0 + u * BK;
                if (!WIDE && u > 0 && kk >= K) break;  // K == BK
                step(u & 1, WIDE && u == 0, kk + 2 * BK);
                if (WIDE && u == UNR - 1) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < TJ; ++j)
#pragma unroll
                            for (int v = 0; v < 4; ++v) wide[WIDE ? i : 0][WIDE ? j : 0][v] += (double)acc[i][j][v];
                }
                __syncthreads();
            }
        }

        TO_GLOBAL float* C = gptr<float>(d->c) + batch * d->cb;
        const float sab = sa * sb;
        const bool vec_out = TR ? d->ldcn == 1 : d->ldcm == 1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                float out[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    out[v] = (WIDE ? (float)wide[WIDE ? i : 0][WIDE ? j : 0][v] : acc[i][j][v]) * sab;
                    if (SCALED) mx = fmaxf(mx, fabsf(out[v]));
                }
                const int64_t mb = tm * BM + wm * 64 + i * 16, nb = tn * BN + wn * 16 * TJ + j * 16;
                if (vec_out) {
                    const int64_t m = TR ? mb + (lane & 15) : mb + Mfma<float>::row(lane, 0);
                    const int64_t n = TR ? nb + Mfma<float>::row(lane, 0) : nb + (lane & 15);
                    *reinterpret_cast<TO_GLOBAL L*>(C + m * d->ldcm + n * d->ldcn) = L{out[0], out[1], out[2], out[3]};
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int64_t m = TR ? mb + (lane & 15) : mb + Mfma<float>::row(lane, v);
                        const int64_t n = TR ? nb + Mfma<float>::row(lane, v) : nb + (lane & 15);
                        C[m * d->ldcm + n * d->ldcn] = out[v];
                    }
                }
            }
        }
        if (SCALED && tid == 0 && tm == 0 && tn == 0) note_exp(d->cs, batch, ea + eb);
    }
    if (SCALED) publish_max_on_change(d->cs, mx_batch, -1, mx, lane);
}

// one block per CU, LDS double-buffered
template <bool SCALED, bool WIDE, bool TR, int BM>
__global__ __launch_bounds__(64 * x6_waves<WIDE>(), 1) void k_gemm_x6(const to_gemm_desc* __restrict__ descs, int ndesc, const int32_t* __restrict__ bmap) {
    gemm_x6_body<SCALED, WIDE, TR, BM>(descs, ndesc, bmap);
}

// ---------------------------------------------------------------------------------
// host dispatch

int grid_ok(const to_launch& L) {
    if (L.grid <= 0 || L.lds < 0 || L.lds > 160 * 1024) return 0;
    return 1;
}

template <typename T, bool SCALED>
int launch_one(const to_launch& L, hipStream_t st) {
    const dim3 grid(L.grid), block(kBlock);
    // descriptor-table launches: optional block -> descriptor map (int32 per block)
    const int32_t* bmap = reinterpret_cast<const int32_t*>(L.aux1);
    switch (L.op) {
        case TO_OP_LEAF_GATHER:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL((k_leaf_gather<T, SCALED>), grid, block, 0, st,
                               reinterpret_cast<const to_leaf_desc*>(L.descs), L.count, bmap,
                               reinterpret_cast<const int64_t*>(L.aux3));
            break;
        case TO_OP_JOIN_STREAM: {
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            const to_join_desc* d = reinterpret_cast<const to_join_desc*>(L.descs);
            switch (L.variant) {
#define TO_STREAM_CASE(YC)                                                                         \
    case YC:                                                                                       \
        hipLaunchKernelGGL((k_join_stream<T, SCALED, YC>), grid, block, L.lds, st, d, L.count, bmap);     \
        break;                                                                                     \
    case TO_STREAM_VEC | YC:                                                                       \
        hipLaunchKernelGGL((k_join_stream_vec<T, SCALED, YC>), grid, block, L.lds, st, d, L.count, bmap); \
        break;
                TO_STREAM_CASE(1)
                TO_STREAM_CASE(2)
                TO_STREAM_CASE(4)
                TO_STREAM_CASE(8)
                TO_STREAM_CASE(16)
#undef TO_STREAM_CASE
                case TO_STREAM_ROWS:
                    hipLaunchKernelGGL((k_join_rows<T, SCALED>), grid, block, 0, st, d, L.count, bmap);
                    break;
                case TO_STREAM_ANY:
                    hipLaunchKernelGGL((k_join_stream_any<T, SCALED>), grid, block, L.lds, st, d, L.count, bmap);
                    break;
                case TO_STREAM_WAVE:
                    if ((int64_t)L.grid * (kBlock / 64) < L.count) {
                        set_err("stream join: a wave launch needs one wave per descriptor");
                        return TO_ERR_INVALID;
                    }
                    hipLaunchKernelGGL((k_join_wave<T, SCALED>), grid, block, 0, st, d, L.count);
                    break;
                default:
                    set_err("stream join: unsupported y chunk");
                    return TO_ERR_INVALID;
            }
            break;
        }
        case TO_OP_JOIN_SPLITK:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL((k_join_splitk<T, SCALED>), grid, block, 0, st,
                               reinterpret_cast<const to_join_desc*>(L.descs), L.count, bmap);
            break;
        case TO_OP_JOIN_FINISH:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            if (L.variant == 1)
                hipLaunchKernelGGL((k_join_finish_wide<T, SCALED>), grid, block, 0, st,
                                   reinterpret_cast<const to_join_desc*>(L.descs), L.count, bmap);
            else
                hipLaunchKernelGGL((k_join_finish<T, SCALED>), grid, block, 0, st,
                                   reinterpret_cast<const to_join_desc*>(L.descs), L.count, bmap);
            break;
        case TO_OP_SCALE_CLEAR:
            if (!L.aux0 || L.iaux0 <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL(k_scale_clear, grid, block, 0, st, reinterpret_cast<to_scale_rec*>(L.aux0),
                               L.iaux0);
            break;
        case TO_OP_ROOT_REDUCE:
            if (!L.aux0 || !L.aux2 || !L.aux3 || L.iaux0 <= 0) return TO_ERR_INVALID;
            if (SCALED && !L.aux1) return TO_ERR_INVALID;
            hipLaunchKernelGGL((k_root_reduce<T, SCALED>), dim3(1), block, 0, st,
                               reinterpret_cast<const T*>(L.aux0),
                               reinterpret_cast<const to_scale_rec*>(L.aux1),
                               reinterpret_cast<to_accum*>(L.aux2), reinterpret_cast<const int64_t*>(L.aux3),
                               (int)L.iaux0);
            break;
        case TO_OP_WAVE_ADVANCE:
            if (!L.aux3) return TO_ERR_INVALID;
            hipLaunchKernelGGL(k_wave_advance, dim3(1), dim3(64), 0, st, reinterpret_cast<int64_t*>(L.aux3),
                               L.iaux0);
            break;
        case TO_OP_PERMUTE:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            if (L.variant == 1)
                hipLaunchKernelGGL((k_permute_vec_desc<T, kPermVecBits<T>>), grid, block, L.lds, st,
                                   reinterpret_cast<const to_perm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 3)
                hipLaunchKernelGGL((k_permute_vec_desc<T, kPermVecBits<T> + 1>), grid, block, L.lds, st,
                                   reinterpret_cast<const to_perm_desc*>(L.descs), L.count, bmap);
            else
                hipLaunchKernelGGL((k_permute_desc<T>), grid, block, L.lds, st,
                                   reinterpret_cast<const to_perm_desc*>(L.descs), L.count, bmap);
            break;
        case TO_OP_GEMM:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            if (L.variant == 1)
                hipLaunchKernelGGL((k_gemm128<T, SCALED, false, false>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 3)
                hipLaunchKernelGGL((k_gemm128<T, SCALED, true, false>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 5)
                hipLaunchKernelGGL((k_gemm128<T, SCALED, false, true>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 7)
                hipLaunchKernelGGL((k_gemm128<T, SCALED, true, true>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 9)
                hipLaunchKernelGGL((k_gemm128_db<T, SCALED, false, false>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 11)
                hipLaunchKernelGGL((k_gemm128_db<T, SCALED, true, false>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 13)
                hipLaunchKernelGGL((k_gemm128_db<T, SCALED, false, true>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 15)
                hipLaunchKernelGGL((k_gemm128_db<T, SCALED, true, true>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant & 16) {
                if constexpr (sizeof(T) == 4) {
                    const to_gemm_desc* gd = reinterpret_cast<const to_gemm_desc*>(L.descs);
                    switch (L.variant) {
#define TO_X6_CASE(V, W, TRN, BMT)                                                                              \
    case V:                                                                                                     \
        hipLaunchKernelGGL((k_gemm_x6<SCALED, W, TRN, BMT>), grid, dim3(64 * x6_waves<W>()), 0, st, gd, L.count, bmap); \
        break;
                        TO_X6_CASE(25, false, false, 128)
                        TO_X6_CASE(27, true, false, 128)
                        TO_X6_CASE(29, false, true, 128)
                        TO_X6_CASE(31, true, true, 128)
                        TO_X6_CASE(57, false, false, 256)
                        TO_X6_CASE(61, false, true, 256)
#undef TO_X6_CASE
                        default:  // the split product runs at one block per CU (bit 3)
                            set_err("gemm: unknown bf16 split-product variant");
                            return TO_ERR_INVALID;
                    }
                } else {
                    set_err("gemm: the bf16 split product takes float operands");
                    return TO_ERR_INVALID;
                }
            }
            else if (L.variant == 0)
                hipLaunchKernelGGL((k_gemm<T, SCALED, false>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else if (L.variant == 2)
                hipLaunchKernelGGL((k_gemm<T, SCALED, true>), grid, block, 0, st,
                                   reinterpret_cast<const to_gemm_desc*>(L.descs), L.count, bmap);
            else {
                set_err("gemm: unknown tile variant");
                return TO_ERR_INVALID;
            }
            break;
        case TO_OP_ROOT_STORE:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL((k_root_store<T, SCALED>), grid, block, 0, st,
                               reinterpret_cast<const to_root_desc*>(L.descs), L.count);
            break;
        case TO_OP_LEAF_BUILD:
            if (!L.descs || L.count <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL((k_leaf_build<T>), grid, block, 0, st,
                               reinterpret_cast<const to_leaf_build_desc*>(L.descs), L.count, bmap);
            break;
        case TO_OP_BATCH_BEGIN:
            if (!L.aux2 || L.iaux0 <= 0) return TO_ERR_INVALID;
            hipLaunchKernelGGL(k_batch_begin, grid, block, 0, st, reinterpret_cast<to_accum*>(L.aux2), L.iaux0,
                               reinterpret_cast<int64_t*>(L.aux3));
            break;
        case TO_OP_STEP_BEGIN:
            if (!L.aux2 || !L.aux3) return TO_ERR_INVALID;
            hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, st, reinterpret_cast<to_accum*>(L.aux2),
                               reinterpret_cast<int64_t*>(L.aux3));
            break;
        default:
            set_err("unknown launch op");
            return TO_ERR_INVALID;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return TO_OK;
}

int run_records(const to_launch* launches, int count, int entry_type, hipStream_t st) {
    if (!launches || count < 0) {
        set_err("null launch list");
        return TO_ERR_INVALID;
    }
    if (entry_type != TO_F64 && entry_type != TO_F32S) {
        set_err("unknown entry type");
        return TO_ERR_INVALID;
    }
    for (int i = 0; i < count; ++i) {
        const to_launch& L = launches[i];
        if (!grid_ok(L)) {
            set_err("bad grid or LDS size in launch record");
            return TO_ERR_INVALID;
        }
        const int rc = entry_type == TO_F64 ? launch_one<double, false>(L, st) : launch_one<float, true>(L, st);
        if (rc != TO_OK) {
            if (rc == TO_ERR_INVALID && g_err[0] == 0) set_err("invalid launch record");
            return rc;
        }
    }
    return TO_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------
// C ABI

extern "C" {

int to_version(void) { return TO_ABI_VERSION; }

int to_struct_sizes(int64_t* join_desc, int64_t* leaf_desc, int64_t* scale_rec, int64_t* accum,
                    int64_t* launch, int64_t* perm_desc, int64_t* gemm_desc, int64_t* root_desc,
                    int64_t* leaf_build_desc) {
    if (leaf_build_desc) *leaf_build_desc = sizeof(to_leaf_build_desc);
    if (root_desc) *root_desc = sizeof(to_root_desc);
    if (perm_desc) *perm_desc = sizeof(to_perm_desc);
    if (gemm_desc) *gemm_desc = sizeof(to_gemm_desc);
    if (join_desc) *join_desc = sizeof(to_join_desc);
    if (leaf_desc) *leaf_desc = sizeof(to_leaf_desc);
    if (scale_rec) *scale_rec = sizeof(to_scale_rec);
    if (accum) *accum = sizeof(to_accum);
    if (launch) *launch = sizeof(to_launch);
    return TO_OK;
}

const char* to_last_error(void) { return g_err; }

int to_device_count(int* count) {
    if (!count) return TO_ERR_INVALID;
    const hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    return TO_OK;
}

int to_set_device(int device) {
    const hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? TO_OK : hip_fail(e, "hipSetDevice");
}

int to_run(const to_launch* launches, int count, int entry_type, void* stream) {
    g_err[0] = 0;
    return run_records(launches, count, entry_type, reinterpret_cast<hipStream_t>(stream));
}

int to_run_timed(const to_launch* launches, int count, int entry_type, void* stream, float* ms) {
    g_err[0] = 0;
    if (!launches || count <= 0 || !ms) {
        set_err("to_run_timed: invalid argument");
        return TO_ERR_INVALID;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t* ev = new hipEvent_t[count + 1];
    int created = 0, rc = TO_OK;
    for (; created <= count; ++created) {
        const hipError_t e = hipEventCreate(&ev[created]);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipEventCreate");
            break;
        }
    }
    if (rc == TO_OK) {
        for (int i = 0; i < count && rc == TO_OK; ++i) {
            hipEventRecord(ev[i], st);
            rc = run_records(launches + i, 1, entry_type, st);
        }
        hipEventRecord(ev[count], st);
        const hipError_t e = hipEventSynchronize(ev[count]);
        if (rc == TO_OK && e != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
        for (int i = 0; i < count && rc == TO_OK; ++i) {
            float t = 0.0f;
            hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            ms[i] = t;
        }
    }
    for (int i = 0; i < created; ++i) hipEventDestroy(ev[i]);
    delete[] ev;
    return rc;
}

int to_graph_create(const to_launch* launches, int count, int entry_type, void* stream, void** graph_exec) {
    g_err[0] = 0;
    if (!graph_exec || !stream) {
        set_err("graph capture needs a non-default stream");
        return TO_ERR_INVALID;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!launches || count < 0) {
        set_err("null launch list");
        return TO_ERR_INVALID;
    }
    bool branched = false;
    for (int i = 0; i < count; ++i) branched |= (launches[i].flags & TO_LAUNCH_SIDE) != 0;
    // the side branch: a second stream that joins the capture through events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    auto release = [&]() {
        if (fork) hipEventDestroy(fork);
        if (join) hipEventDestroy(join);
        if (side) hipStreamDestroy(side);
    };
    hipError_t e = hipSuccess;
    if (branched) {
        e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "side stream");
        }
    }
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "hipStreamBeginCapture");
    }
    auto fork_side = [&]() {  // the side stream continues from the main stream's point
        hipError_t r = hipEventRecord(fork, st);
        return r == hipSuccess ? hipStreamWaitEvent(side, fork, 0) : r;
    };
    auto join_side = [&]() {  // the main stream waits for everything on the side stream
        hipError_t r = hipEventRecord(join, side);
        return r == hipSuccess ? hipStreamWaitEvent(st, join, 0) : r;
    };
    int rc = TO_OK;
    if (branched) {
        const hipError_t r = fork_side();
        if (r != hipSuccess) rc = hip_fail(r, "fork");
    }
    for (int i = 0; i < count && rc == TO_OK; ++i) {
        const to_launch& L = launches[i];
        if (branched && (L.flags & TO_LAUNCH_JOIN)) {
            hipError_t r = join_side();
            if (r == hipSuccess) r = fork_side();
            if (r != hipSuccess) {
                rc = hip_fail(r, "join");
                break;
            }
        }
        rc = run_records(&L, 1, entry_type, branched && (L.flags & TO_LAUNCH_SIDE) ? side : st);
    }
    if (branched) {
        const hipError_t r = join_side();  // the capture must end joined
        if (rc == TO_OK && r != hipSuccess) rc = hip_fail(r, "join");
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    release();
    if (rc != TO_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = reinterpret_cast<void*>(exec);
    return TO_OK;
}

int to_graph_launch(void* graph_exec, void* stream) {
    g_err[0] = 0;
    if (!graph_exec) return TO_ERR_INVALID;
    const hipError_t e =
        hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? TO_OK : hip_fail(e, "hipGraphLaunch");
}

int to_graph_destroy(void* graph_exec) {
    if (!graph_exec) return TO_OK;
    const hipError_t e = hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec));
    return e == hipSuccess ? TO_OK : hip_fail(e, "hipGraphExecDestroy");
}

int to_permute(const void* in, void* out, int rank, const int32_t* perm, int elem_bytes, int64_t batch,
               void* stream) {
    g_err[0] = 0;
    if (!in || !out || !perm || rank < 0 || rank > 31 || batch <= 0 || batch > INT32_MAX ||
        (elem_bytes != 4 && elem_bytes != 8)) {
        set_err("to_permute: invalid argument");
        return TO_ERR_INVALID;
    }
    // opos[p]: output bit position of input bit position p
    int opos[32], seen[32] = {0};
    for (int j = 0; j < rank; ++j) {
        const int a = perm[j];
        if (a < 0 || a >= rank || seen[a]) {
            set_err("to_permute: perm is not a permutation");
            return TO_ERR_INVALID;
        }
        seen[a] = 1;
        opos[rank - 1 - a] = rank - 1 - j;
    }
    int ipos_of_out[32];
    for (int p = 0; p < rank; ++p) ipos_of_out[opos[p]] = p;
    // grow m while the tile union stays within the tile bits; a tensor of at least that
    // many bits gets exactly that many (extra low input bits), which the vectorised
    // kernel needs
    const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    // 16-byte chunk tiles of 64 KB (2^14 floats / 2^13 doubles) when the tensor has that
    // many bits, else 32 KB; the scalar kernel takes <= 12 bits
    const int small_bits = elem_bytes == 4 ? 13 : 12;
    const int tile_bits = !aligned ? 12 : (rank > small_bits ? small_bits + 1 : small_bits);
    int m = 0, inU[32] = {0};
    for (;;) {
        if (m >= rank) break;
        int cand[32];
        memcpy(cand, inU, sizeof(cand));
        cand[m] = 1;
        cand[ipos_of_out[m]] = 1;
        int cnt = 0;
        for (int p = 0; p < rank; ++p) cnt += cand[p];
        if (cnt > tile_bits) break;
        memcpy(inU, cand, sizeof(cand));
        ++m;
    }
    {
        int cnt = 0;
        for (int q = 0; q < rank; ++q) cnt += inU[q];
        for (int q = 0; q < rank && cnt < tile_bits; ++q)
            if (!inU[q]) {
                inU[q] = 1;
                ++cnt;
            }
    }
    to_perm_desc p;
    memset(&p, 0, sizeof(p));
    p.in = reinterpret_cast<uint64_t>(in);
    p.out = reinterpret_cast<uint64_t>(out);
    p.inb = p.outb = 1ll << rank;
    p.rank = rank;
    p.nbatch = (int32_t)batch;
    int nu = 0;
    for (int q = 0; q < rank; ++q)
        if (inU[q]) p.u_in[nu++] = (int8_t)q;
    p.nu = nu;
    int v = 0;
    for (int q = 0; q < rank; ++q) {
        const int ip = ipos_of_out[q];
        if (inU[ip]) {
            p.v_out[v] = (int8_t)q;
            for (int j = 0; j < nu; ++j)
                if (p.u_in[j] == ip) p.v2u[v] = (int8_t)j;
            ++v;
        }
    }
    int nr = 0;
    for (int q = 0; q < rank; ++q) {
        if (!inU[q]) {
            p.rest_in[nr] = (int8_t)q;
            p.rest_out[nr] = (int8_t)opos[q];
            ++nr;
        }
    }
    p.nrest = nr;
    const int64_t ntiles = batch << nr;
    const bool vec = aligned && nu == tile_bits && nu >= small_bits;
    const int64_t cap = vec ? 1024 : 4096;
    const int grid = (int)(ntiles < cap ? ntiles : cap);
    p.block0 = 0;
    p.nblocks = grid;
    const size_t lds = (((size_t)elem_bytes << nu) + 15) / 16 * 16 + 2 * 256 * sizeof(int64_t);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const bool big = nu > small_bits;
    if (vec && elem_bytes == 8 && big)
        hipLaunchKernelGGL((k_permute_vec<double, kPermVecBits<double> + 1>), dim3(grid), dim3(kBlock), lds, st, p);
    else if (vec && elem_bytes == 8)
        hipLaunchKernelGGL((k_permute_vec<double, kPermVecBits<double>>), dim3(grid), dim3(kBlock), lds, st, p);
    else if (vec && big)
        hipLaunchKernelGGL((k_permute_vec<float, kPermVecBits<float> + 1>), dim3(grid), dim3(kBlock), lds, st, p);
    else if (vec)
        hipLaunchKernelGGL((k_permute_vec<float, kPermVecBits<float>>), dim3(grid), dim3(kBlock), lds, st, p);
    else if (elem_bytes == 8)
        hipLaunchKernelGGL(k_permute<double>, dim3(grid), dim3(kBlock), lds, st, p);
    else
        hipLaunchKernelGGL(k_permute<float>, dim3(grid), dim3(kBlock), lds, st, p);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? TO_OK : hip_fail(e, "to_permute launch");
}

}  // extern "C"

namespace {

// One tensordot as one join: the bit geometry, the kernel form (streamed join, or split-K
// with double partial sums and a finishing pass) and the scratch it needs.
struct DotPlan {
    to_join_desc d;
    bool x_is_left;  // a is the streamed operand X
    bool splitk;
    int yc, lds, grid, fin_grid;
    bool fin_wide;
    int64_t partial_bytes;
};

constexpr int kDotMaxRank = 30;
constexpr int kDotLdsLimit = 64 * 1024;
constexpr int kDotMaxBlocks = 4096;

// numpy.tensordot(a, b, (axes_a, axes_b)) of all-dimension-2 tensors: output axes are a's
// unshared axes in order, then b's (numpy/_core/numeric.py:1161-1178); axis i of a rank-r
// tensor is bit r-1-i of its flat index.  The operand with more free bits is streamed (X).
int plan_tensordot(int rank_a, int rank_b, const int32_t* axes_a, const int32_t* axes_b, int nshared,
                   int entry_type, DotPlan& P) {
    if (entry_type != TO_F64 && entry_type != TO_F32S) {
        set_err("tensordot: unknown entry type");
        return TO_ERR_INVALID;
    }
    if (rank_a < 0 || rank_b < 0 || rank_a > kDotMaxRank || rank_b > kDotMaxRank || nshared < 0 ||
        nshared > rank_a || nshared > rank_b || (nshared > 0 && (!axes_a || !axes_b))) {
        set_err("tensordot: invalid ranks or shared-axis count");
        return TO_ERR_INVALID;
    }
    bool in_a[32] = {false}, in_b[32] = {false};
    for (int t = 0; t < nshared; ++t) {
        const int i = axes_a[t], j = axes_b[t];
        if (i < 0 || i >= rank_a || j < 0 || j >= rank_b || in_a[i] || in_b[j]) {
            set_err("tensordot: axes out of range or repeated");
            return TO_ERR_INVALID;
        }
        in_a[i] = in_b[j] = true;
    }
    const int s = nshared, na = rank_a - s, nb = rank_b - s;
    if (na + nb > kDotMaxRank) {
        set_err("tensordot: output rank above 30");
        return TO_ERR_INVALID;
    }
    int fa[32], fb[32], k = 0;
    for (int i = 0; i < rank_a; ++i)
        if (!in_a[i]) fa[k++] = i;
    k = 0;
    for (int i = 0; i < rank_b; ++i)
        if (!in_b[i]) fb[k++] = i;
    int8_t a_free[32], b_free[32], a_sh[32], b_sh[32];
    for (int t = 0; t < na; ++t) a_free[t] = (int8_t)(rank_a - 1 - fa[na - 1 - t]);
    for (int t = 0; t < nb; ++t) b_free[t] = (int8_t)(rank_b - 1 - fb[nb - 1 - t]);
    for (int t = 0; t < s; ++t) {
        a_sh[t] = (int8_t)(rank_a - 1 - axes_a[s - 1 - t]);
        b_sh[t] = (int8_t)(rank_b - 1 - axes_b[s - 1 - t]);
    }
    const bool x_is_left = na > nb || (na == nb && rank_a >= rank_b);
    memset(&P, 0, sizeof(P));
    to_join_desc& d = P.d;
    const int fx = x_is_left ? na : nb, fy = x_is_left ? nb : na;
    d.fx = fx;
    d.fy = fy;
    d.s = s;
    d.shf = x_is_left ? nb : 0;
    d.shy = x_is_left ? 0 : nb;
    d.nbatch = 1;
    memcpy(d.xf_pos, x_is_left ? a_free : b_free, (size_t)fx);
    memcpy(d.yf_pos, x_is_left ? b_free : a_free, (size_t)fy);
    memcpy(d.xs_pos, x_is_left ? a_sh : b_sh, (size_t)s);
    memcpy(d.ys_pos, x_is_left ? b_sh : a_sh, (size_t)s);
    P.x_is_left = x_is_left;
    const int64_t K = 1ll << s;
    if (s <= 12) {
        int yc = fy >= 4 ? 16 : 1 << fy;
        int64_t lds = 8 * K;
        if (fx >= 8) {
            // the Y slab is held in LDS as the double accumulator type
            while (yc > 1 && 8 * K + yc * K * 8 > kDotLdsLimit) yc >>= 1;
            lds = 8 * K + yc * K * 8;
        }
        const int64_t nych = (1ll << fy) / yc;
        if (lds <= kDotLdsLimit && !(s >= 8 && (nych << fx) < 32768)) {
            const int64_t ntiles = fx >= 8 ? (nych << (fx - 8)) : ((1ll << fx) + kBlock - 1) / kBlock * nych;
            P.splitk = false;
            P.yc = yc;
            P.lds = (int)lds;
            P.grid = (int)(ntiles < kDotMaxBlocks ? ntiles : kDotMaxBlocks);
            int l = 0;
            while ((1ll << l) < nych) ++l;
            d.nych = (int32_t)nych;
            d.lnych = l;
            d.nsplit = 1;
            d.nblocks = P.grid;
            return TO_OK;
        }
    }
    // split-K: block (o, sp) sums its share of the 2^s shared indices of output o
    const int64_t nout = 1ll << (fx + fy);
    int lsplit = 0;
    if (s >= 8) {
        const int64_t room = nout < 4096 ? 4096 / nout : 1;
        while ((2ll << lsplit) <= room && lsplit < s - 8) ++lsplit;
    }
    const int64_t nsplit = 1ll << lsplit;
    if (nout * nsplit > 0x7fffffffll) {
        set_err("tensordot: too many outputs for the split-K form");
        return TO_ERR_INVALID;
    }
    P.splitk = true;
    P.grid = (int)(nout * nsplit);
    P.fin_wide = nsplit >= 32;
    const int64_t fin = P.fin_wide ? nout : (nout + kBlock - 1) / kBlock;
    P.fin_grid = (int)(fin < kDotMaxBlocks ? fin : kDotMaxBlocks);
    P.partial_bytes = nout * nsplit * (int64_t)sizeof(double);
    d.nsplit = (int32_t)nsplit;
    d.nblocks = P.grid;
    return TO_OK;
}

// Launch a planned tensordot: clear the output's scale record (TO_F32S), then the join.
template <typename T, bool SCALED>
int launch_tensordot(DotPlan& P, hipStream_t st) {
    if (SCALED) hipLaunchKernelGGL(k_scale_clear, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<to_scale_rec*>(P.d.cs), (int64_t)1);
    const dim3 block(kBlock);
    if (!P.splitk) {
        switch (P.yc) {
#define TO_DOT_CASE(YC)                                                                                   \
    case YC:                                                                                              \
        hipLaunchKernelGGL((k_join_stream_one<T, SCALED, YC>), dim3(P.grid), block, P.lds, st, P.d);     \
        break;
            TO_DOT_CASE(1)
            TO_DOT_CASE(2)
            TO_DOT_CASE(4)
            TO_DOT_CASE(8)
            TO_DOT_CASE(16)
#undef TO_DOT_CASE
            default:
                set_err("tensordot: bad y chunk");
                return TO_ERR_INVALID;
        }
    } else {
        hipLaunchKernelGGL((k_join_splitk_one<T, SCALED>), dim3(P.grid), block, 0, st, P.d);
        to_join_desc fin = P.d;
        fin.nblocks = P.fin_grid;  // the finishing kernels stride over outputs by their grid
        if (P.fin_wide)
            hipLaunchKernelGGL((k_join_finish_wide_one<T, SCALED>), dim3(P.fin_grid), block, 0, st, fin);
        else
            hipLaunchKernelGGL((k_join_finish_one<T, SCALED>), dim3(P.fin_grid), block, 0, st, fin);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? TO_OK : hip_fail(e, "tensordot launch");
}

int run_tensordot(DotPlan& P, const void* a, const void* b, void* c, int entry_type, const void* a_rec,
                  const void* b_rec, void* c_rec, void* partial, hipStream_t st) {
    if (!a || !b || !c || (entry_type == TO_F32S && (!a_rec || !b_rec || !c_rec))) {
        set_err("tensordot: null tensor or scale record");
        return TO_ERR_INVALID;
    }
    const bool x_is_left = P.x_is_left;
    P.d.x = reinterpret_cast<uint64_t>(x_is_left ? a : b);
    P.d.y = reinterpret_cast<uint64_t>(x_is_left ? b : a);
    P.d.c = reinterpret_cast<uint64_t>(c);
    if (entry_type == TO_F32S) {
        P.d.xs = reinterpret_cast<uint64_t>(x_is_left ? a_rec : b_rec);
        P.d.ys = reinterpret_cast<uint64_t>(x_is_left ? b_rec : a_rec);
        P.d.cs = reinterpret_cast<uint64_t>(c_rec);
    }
    P.d.partial = reinterpret_cast<uint64_t>(partial);
    return entry_type == TO_F64 ? launch_tensordot<double, false>(P, st) : launch_tensordot<float, true>(P, st);
}

// Scratch of to_tensordot_noalloc: the split-K partial sums (doubles) of one call, in a
// file-scope device array (every GPU the library runs on holds one copy until the process
// ends).  Host code never passes the array's name to a kernel: it is a host-side stand-in,
// not a device address; each device's address comes from hipGetSymbolAddress, once per
// device, with that device current.
constexpr int64_t kNoAllocScratchBytes = 8ll << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};

int scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_err("scratch: device index above 63");
        return TO_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TO_OK;
}

}  // namespace

extern "C" {

int to_tensordot_noalloc(const void* a, const void* b, void* c, int rank_a, int rank_b, const int32_t* axes_a,
                         const int32_t* axes_b, int nshared, int entry_type, const void* a_rec, const void* b_rec,
                         void* c_rec, void* stream) {
    g_err[0] = 0;
    void* ptr = nullptr;
    int rc = scratch_address(&ptr);
    if (rc != TO_OK) return rc;
    DotPlan P;
    rc = plan_tensordot(rank_a, rank_b, axes_a, axes_b, nshared, entry_type, P);
    if (rc != TO_OK) return rc;
    if (P.partial_bytes > kNoAllocScratchBytes) {
        snprintf(g_err, sizeof(g_err),
                 "to_tensordot_noalloc: needs %lld bytes of partial sums; its scratch holds %lld (use to_tensordot)",
                 (long long)P.partial_bytes, (long long)kNoAllocScratchBytes);
        return TO_ERR_INVALID;
    }
    return run_tensordot(P, a, b, c, entry_type, a_rec, b_rec, c_rec, ptr, reinterpret_cast<hipStream_t>(stream));
}

int to_tensordot_workspace(int rank_a, int rank_b, const int32_t* axes_a, const int32_t* axes_b, int nshared,
                           int entry_type, int64_t* bytes) {
    g_err[0] = 0;
    if (!bytes) {
        set_err("to_tensordot_workspace: null output");
        return TO_ERR_INVALID;
    }
    DotPlan P;
    const int rc = plan_tensordot(rank_a, rank_b, axes_a, axes_b, nshared, entry_type, P);
    *bytes = rc == TO_OK ? P.partial_bytes : 0;
    return rc;
}

int to_tensordot(const void* a, const void* b, void* c, int rank_a, int rank_b, const int32_t* axes_a,
                 const int32_t* axes_b, int nshared, int entry_type, const void* a_rec, const void* b_rec,
                 void* c_rec, void* workspace, int64_t workspace_bytes, void* stream) {
    g_err[0] = 0;
    DotPlan P;
    int rc = plan_tensordot(rank_a, rank_b, axes_a, axes_b, nshared, entry_type, P);
    if (rc != TO_OK) return rc;
    if (P.partial_bytes > 0 &&
        (!workspace || workspace_bytes < P.partial_bytes || (reinterpret_cast<uintptr_t>(workspace) & 7))) {
        snprintf(g_err, sizeof(g_err), "to_tensordot: needs an 8-byte aligned workspace of %lld bytes",
                 (long long)P.partial_bytes);
        return TO_ERR_INVALID;
    }
    return run_tensordot(P, a, b, c, entry_type, a_rec, b_rec, c_rec, workspace,
                         reinterpret_cast<hipStream_t>(stream));
}

int to_wall_clock_khz(int* khz) {
    if (!khz) return TO_ERR_INVALID;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev);
    return e == hipSuccess ? TO_OK : hip_fail(e, "hipDeviceGetAttribute(WallClockRate)");
}

int to_synchronize(void* stream) {
    const hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? TO_OK : hip_fail(e, "hipStreamSynchronize");
}

}  // extern "C"
