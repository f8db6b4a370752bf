// Batched negacyclic NTT / INTT over 60-bit primes for gfx950.
//
// Computes the same transform as the reference's two-kernel radix-8 NTT
// (include/ntt_60bit.cuh:414-823, dispatch include/BFVcontext.cuh:1513-1552):
// Cooley-Tukey forward with the bit-reversed psi table (stage with m groups uses
// psi[m + group]), Gentleman-Sande inverse with the bit-reversed psi^-1 table
// whose entry 1 carries n^-1, outputs canonical in [0, q).
//
// Design (MI355X): N = N1 * N2 four-step split into two passes, each a kernel
// that stages a tile of 2^11-2^13 elements (16-64 KiB, the plan's choice) in
// LDS and runs radix-16 rounds in registers:
//   column pass  -- C adjacent columns of stride N/N1 (coalesced C*8-byte rows),
//                   global stages 0 .. log2(N1)-1
//   row pass     -- C contiguous chunks of N2 elements, remaining stages
// Forward = column then row; inverse = row then column.  Values stay lazy in
// [0, 8q) (forward) / [0, 4q) (inverse) between stages (moduli are below 2^60,
// so every sum stays below 2^63) and are reduced once on the final store.  The
// twiddle products use an under-estimated Shoup quotient (one 32-bit high
// product and two carries fewer, result in [0, 4q)), and conditional
// subtractions select on the sign word with v_bfi_b32, so no carry or compare
// result passes through an SGPR.  Twiddles are (w, Shoup w') pairs read as one 16-byte load,
// one load per distinct twiddle per thread.  The grid is ordered limb-slowest so
// that every workgroup in flight shares one modulus' tables in L2.
//
// Plans: the split point (column pass length 2^LOG_S, the row pass takes the
// rest of log N), the tile of each pass (2^11..2^13 elements per workgroup,
// 16 values per thread) and the XCD-ordered grid of each pass are template /
// launch parameters; every plan computes the same bits.  Passes of length 2^9
// run a radix-32 round and a radix-16 round.  A context keeps one plan per
// (direction, transform-count bucket); kbfv_ntt_autotune times the candidates
// on the caller's stream and stores the fastest.
#include <algorithm>
#include <type_traits>

#include "kbfv_internal.h"

namespace {

constexpr int LOG_R = 4;  // 16 values per thread

// a pass of 2^log_s-point sub-transforms over tiles of 2^log_t elements: the
// kernels instantiated for it and the host's plan validation agree on this
__host__ __device__ constexpr int round0_log(int log_s) { return log_s <= 8 ? (log_s < 4 ? log_s : 4) : 5; }
__host__ __device__ constexpr bool pass_ok(int log_n, int log_s, int log_t, bool column) {
    return log_s >= 5 && log_s <= 9 && log_t >= 11 && log_t <= 13 && log_t >= log_s && log_t <= log_n &&
           (!column || log_t - log_s >= 3) &&  // column rows of >= 64 bytes
           log_s - round0_log(log_s) >= 1 && log_s - round0_log(log_s) <= LOG_R;
}
__host__ __device__ constexpr int col_log_default(int log_n) { return (log_n + 1) / 2; }

template <int LOG_S, bool COLUMN, int LOG_T>
struct Geo {
    static constexpr int TILE = 1 << LOG_T;    // elements per workgroup
    static constexpr int S = 1 << LOG_S;
    static constexpr int C = TILE / S;         // sub-transforms per workgroup
    static constexpr int T = TILE >> LOG_R;    // threads per workgroup
    static constexpr int ROWP = S + S / 16;    // row pass: one pad word per 16
    static constexpr int WORDS = COLUMN ? TILE : C * ROWP;
    __device__ static __forceinline__ int addr(int c, int k) {
        if constexpr (COLUMN)
            return k * C + c;
        else
            return c * ROWP + k + (k >> 4);
    }
};

// 64-bit value + 32-bit value as one v_mad_u64_u32 (b * 1 + a): no carry
// through an SGPR and no register pair to build for b
__device__ __forceinline__ uint64_t add_u32(uint64_t a, uint32_t b) {
    uint32_t one = 1;
    asm("" : "+s"(one));  // a multiplier the optimizer cannot fold away
    return kb_mad32(b, one, a);
}

// Shoup product x*w mod q with an under-estimated quotient, in [0, 4q) for any
// x < 2^64.  floor(x*ws / 2^64) = x1*s1 + hi(x1*s0) + hi(x0*s1) + e with
// e in {0, 1, 2} (the carries of the low words and hi(x0*s0)); leaving e out
// adds at most 2q to the lazy residue x*w - h*q in [0, 2q).  nq = 2^64 - q.
__device__ __forceinline__ uint64_t shoup4(uint64_t x, uint64_t w, uint64_t ws, uint64_t nq) {
    const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
    const uint32_t s0 = (uint32_t)ws, s1 = (uint32_t)(ws >> 32);
    const uint64_t h = add_u32(kb_mad32(x1, s1, __umulhi(x1, s0)), __umulhi(x0, s1));
    const uint32_t h0 = (uint32_t)h, h1 = (uint32_t)(h >> 32);
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
    const uint32_t n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
    const uint64_t r = kb_mad32(h0, n0, kb_mad32(x0, w0, 0));
    const uint32_t cross = kb_mul_lo32(x0, w1) + kb_mul_lo32(x1, w0) + kb_mul_lo32(h0, n1) +
                           kb_mul_lo32(h1, n0);
    return r + ((uint64_t)cross << 32);
}

// a >= m ? a - m : a for a, m < 2^63, given nm = 2^64 - m: the sign of a - m
// selects word by word (v_bfi_b32), no compare or carry through an SGPR
__device__ __forceinline__ uint64_t csub_sel(uint64_t a, uint64_t nm) {
    const uint64_t t = a + nm;
    const uint32_t msk = (uint32_t)((int32_t)(uint32_t)(t >> 32) >> 31);
    uint32_t lo, hi;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(lo) : "v"(msk), "v"((uint32_t)a), "v"((uint32_t)t));
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(hi) : "v"(msk), "v"((uint32_t)(a >> 32)), "v"((uint32_t)(t >> 32)));
    return ((uint64_t)hi << 32) | lo;
}

// forward values stay in [0, 8q), inverse values in [0, 4q); q < 2^60 keeps
// every sum below 2^63.  nq = 2^64 - q, q4 = 4q, nq4 = 2^64 - 4q.
__device__ __forceinline__ void ct_bfly(uint64_t &a, uint64_t &b, kbfv_tw w, uint64_t nq, uint64_t q4,
                                        uint64_t nq4) {
    a = csub_sel(a, nq4);                        // [0, 4q)
    const uint64_t t = shoup4(b, w.w, w.ws, nq);  // [0, 4q)
    b = a + q4 - t;                              // (0, 8q)
    a = a + t;                                   // [0, 8q)
}

__device__ __forceinline__ void gs_bfly(uint64_t &a, uint64_t &b, kbfv_tw w, uint64_t nq, uint64_t q4,
                                        uint64_t nq4) {
    const uint64_t u = csub_sel(a + b, nq4);               // [0, 4q)
    const uint64_t v = shoup4(a + q4 - b, w.w, w.ws, nq);  // [0, 4q)
    a = u;
    b = v;
}

// The twiddles of the one butterfly group thread threadIdx.x takes in a round
// (rounds with exactly one group per thread), in the order the round consumes
// them: level j of the round, sub-block m -> slot (1 << j) - 1 + m.
template <int LR, int LOG_S, bool COLUMN, int LOG_T>
__device__ __forceinline__ void round_twiddles(kbfv_tw *tws, const int lam0, const kbfv_tw *tab,
                                               const int log_n, const int g0, const int grp0,
                                               const int stride) {
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int S = G::S, C = G::C, RR = 1 << LR;
    const int grp = threadIdx.x;
    int c, u;
    if constexpr (COLUMN) {
        c = grp % C;
        u = grp / C;
    } else {
        u = grp % (S >> LR);
        c = grp / (S >> LR);
    }
    const int Bs = S >> lam0;
    const int d = Bs >> LR;
    const int kbase = (u / d) * Bs + (u % d);
    const int posc = COLUMN ? (grp0 * C + c) : (grp0 * C + c) * S;
    const int pstep = COLUMN ? stride : 1;
#pragma unroll
    for (int j = 0; j < LR; ++j) {
        const int h = RR >> (j + 1);
        const int gst = g0 + lam0 + j;
#pragma unroll
        for (int m = 0; m < (1 << j); ++m) {
            const int pos = posc + (kbase + m * (2 * h) * d) * pstep;
            tws[(1 << j) - 1 + m] = tab[(1 << gst) + (pos >> (log_n - gst))];
        }
    }
}

// One register round of LR levels starting at pass level lam0.  The round reads
// its 16 values from the LDS tile or, when GIN, straight from global memory
// (the rounds whose lanes cover adjacent columns / 16 adjacent elements, i.e.
// 128-byte runs); it writes them back to the tile or, when GOUT, straight to
// global memory with the final reduction of the last pass (FIN).
template <int LR, int LOG_S, bool COLUMN, int LOG_T, bool FWD, bool NINV, bool GIN = false,
          bool GOUT = false, bool FIN = false, bool PRE = false, bool W1 = false, bool DIG = false>
__device__ __forceinline__ void ntt_round(uint64_t *tile, const int lam0, const kbfv_tw *tab,
                                          const int log_n, const int g0, const int grp0,
                                          const int stride, const uint64_t q, const kbfv_tw ninv,
                                          const uint64_t *__restrict__ gin = nullptr,
                                          uint64_t *__restrict__ gout = nullptr,
                                          const kbfv_tw (*pre)[(1 << LR) - 1] = nullptr,
                                          const int dshift = 0, const uint64_t dmask = ~0ULL) {
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int S = G::S, C = G::C, RR = 1 << LR;
    constexpr int NG = C * (S >> LR);
    const uint64_t q4 = q << 2;
    uint64_t nq = 0 - q, nq4 = 0 - q4;
    // opaque to the optimizer, which would otherwise turn x + (2^64 - m) back
    // into a subtraction with its carry through an SGPR
    asm volatile("" : "+s"(nq), "+s"(nq4));
    const int Bs = S >> lam0;
    const int d = Bs >> LR;
    for (int grp = threadIdx.x; grp < NG; grp += G::T) {
        int c, u;
        if constexpr (COLUMN) {
            c = grp % C;
            u = grp / C;
        } else {
            u = grp % (S >> LR);
            c = grp / (S >> LR);
        }
        const int kbase = (u / d) * Bs + (u % d);
        // global position of element k of sub-transform c
        const int posc = COLUMN ? (grp0 * C + c) : (grp0 * C + c) * S;
        const int pstep = COLUMN ? stride : 1;
        uint64_t x[RR];
#pragma unroll
        for (int i = 0; i < RR; ++i) {
            if constexpr (GIN) {
                x[i] = gin[posc + (kbase + i * d) * pstep];
                // DIG: the input is one base-2^w digit of the loaded word
                if constexpr (DIG) x[i] = (x[i] >> dshift) & dmask;
            } else
                x[i] = tile[G::addr(c, kbase + i * d)];
        }
        if constexpr (FWD) {
#pragma unroll
            for (int j = 0; j < LR; ++j) {
                const int h = RR >> (j + 1);
                const int gst = g0 + lam0 + j;
#pragma unroll
                for (int m = 0; m < (1 << j); ++m) {
                    const int i0 = m * (2 * h);
                    const int pos = posc + (kbase + i0 * d) * pstep;
                    kbfv_tw w;
                    if constexpr (PRE)
                        w = (*pre)[(1 << j) - 1 + m];
                    else
                        w = tab[(1 << gst) + (pos >> (log_n - gst))];
#pragma unroll
                    for (int i = i0; i < i0 + h; ++i) ct_bfly(x[i], x[i + h], w, nq, q4, nq4);
                }
            }
        } else {
#pragma unroll
            for (int j = LR - 1; j >= 0; --j) {
                const int h = RR >> (j + 1);
                const int gst = g0 + lam0 + j;
#pragma unroll
                for (int m = 0; m < (1 << j); ++m) {
                    const int i0 = m * (2 * h);
                    const int pos = posc + (kbase + i0 * d) * pstep;
                    kbfv_tw w;
                    if constexpr (PRE)
                        w = (*pre)[(1 << j) - 1 + m];
                    else
                        w = tab[(1 << gst) + (pos >> (log_n - gst))];
                    // W1: stage 0 takes the unscaled entry 1, passed in ninv
                    if (W1 && gst == 0) w = ninv;
#pragma unroll
                    for (int i = i0; i < i0 + h; ++i) {
                        gs_bfly(x[i], x[i + h], w, nq, q4, nq4);
                        if (NINV && gst == 0) x[i] = shoup4(x[i], ninv.w, ninv.ws, nq);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < RR; ++i) {
            if constexpr (GOUT) {
                uint64_t v = x[i];
                if constexpr (FIN) {
                    if constexpr (FWD) v = csub_sel(v, nq4);
                    v = csub_sel(v, 0 - (q << 1));
                    v = csub_sel(v, nq);
                }
                gout[posc + (kbase + i * d) * pstep] = v;
            } else {
                tile[G::addr(c, kbase + i * d)] = x[i];
            }
        }
    }
    if constexpr (!GOUT) __syncthreads();
}

// coalesced global -> LDS copy of the workgroup's tile
template <int LOG_S, bool COLUMN, int LOG_T>
__device__ __forceinline__ void stage_in(uint64_t *tile, const uint64_t *__restrict__ in, int grp0,
                                         int stride) {
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int S = G::S, C = G::C, T = G::T;
#pragma unroll
    for (int j = 0; j < G::TILE / T; ++j) {
        const int e = threadIdx.x + j * T;
        int c, k;
        if constexpr (COLUMN) {
            c = e % C;
            k = e / C;
        } else {
            c = e / S;
            k = e % S;
        }
        const int pos = COLUMN ? (grp0 * C + c + k * stride) : ((grp0 * C + c) * S + k);
        tile[G::addr(c, k)] = in[pos];
    }
}

// one workgroup's tile of one limb: `in` / `out` address the limb, m is its modulus
// DIG (forward column pass only): the input value is (word >> dshift) & dmask
template <int LOG_N, int LOG_S, int LOG_T, bool COLUMN, bool FWD, bool NOSCALE, bool DIG = false>
__device__ __forceinline__ void pass_body(uint64_t *tile, const uint64_t *__restrict__ in,
                                          uint64_t *__restrict__ out, const kbfv_tw *__restrict__ tw,
                                          const kbfv_tw *__restrict__ ninv_tab, const kbfv_tw *__restrict__ w1_tab,
                                          const uint64_t *__restrict__ mods, const int m, const int grp0,
                                          const int dshift = 0, const uint64_t dmask = ~0ULL) {
    static_assert(!DIG || (COLUMN && FWD), "digit input only on the forward column pass");
    constexpr int log_n = LOG_N;
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int S = G::S, C = G::C, T = G::T;
    static_assert(pass_ok(LOG_N, LOG_S, LOG_T, COLUMN), "pass shape outside the plan space");
    constexpr bool LAST = (FWD != COLUMN);  // forward ends with the row pass
    constexpr int n = 1 << log_n;
    const uint64_t q = mods[m];
    const kbfv_tw *tab = tw + ((size_t)m * 2 + (FWD ? 0 : 1)) * n;
    constexpr int stride = n >> LOG_S;
    constexpr int g0 = COLUMN ? 0 : (log_n - LOG_S);

    constexpr int LR0 = round0_log(LOG_S);
    constexpr int LR1 = LOG_S - LR0;
    static_assert(LR1 > 0 && LR1 <= LOG_R, "two register rounds per pass");
    // NOSCALE (inverse column pass): N x INTT, lazy, stage 0 with the
    // unscaled twiddle (passed as w1_tab) and no n^-1 multiply
    constexpr bool NI = (!FWD && LAST && !NOSCALE);
    const kbfv_tw ninv = NI ? ninv_tab[m] : (NOSCALE ? w1_tab[m] : kbfv_tw{0, 0});
    // the forward pass's first register round (stride S/16) and the inverse pass's last
    // touch 128-byte runs in both layouts; the stride-1 rounds of the row pass
    // go through the LDS tile for both their global load and store.
    if constexpr (FWD) {
        ntt_round<LR0, LOG_S, COLUMN, LOG_T, true, false, true, false, false, false, false, DIG>(
            tile, 0, tab, log_n, g0, grp0, stride, q, ninv, in, nullptr, nullptr, dshift, dmask);
        if constexpr (COLUMN) {
            ntt_round<LR1, LOG_S, COLUMN, LOG_T, true, false, false, true, LAST>(tile, LR0, tab, log_n, g0, grp0,
                                                                                 stride, q, ninv, nullptr, out);
            return;
        } else {
            ntt_round<LR1, LOG_S, COLUMN, LOG_T, true, false>(tile, LR0, tab, log_n, g0, grp0, stride, q, ninv);
        }
    } else {
        if constexpr (COLUMN) {
            ntt_round<LR1, LOG_S, COLUMN, LOG_T, false, NI, true>(tile, LR0, tab, log_n, g0, grp0, stride, q,
                                                                  ninv, in);
        } else {
            // one butterfly group per thread (a radix-16 round): its twiddles are
            // loaded with the tile, ahead of the barrier, rather than after it
            constexpr bool PRE = (C * (S >> LR1) == T);
            if constexpr (PRE) {
                kbfv_tw pre[(1 << LR1) - 1];
                round_twiddles<LR1, LOG_S, COLUMN, LOG_T>(pre, LR0, tab, log_n, g0, grp0, stride);
                stage_in<LOG_S, COLUMN, LOG_T>(tile, in, grp0, stride);
                // pinned here so that the loads are not sunk below the barrier
#pragma unroll
                for (int i = 0; i < (1 << LR1) - 1; ++i) asm volatile("" : "+v"(pre[i].w), "+v"(pre[i].ws));
                __syncthreads();
                ntt_round<LR1, LOG_S, COLUMN, LOG_T, false, NI, false, false, false, true>(
                    tile, LR0, tab, log_n, g0, grp0, stride, q, ninv, nullptr, nullptr, &pre);
            } else {
                stage_in<LOG_S, COLUMN, LOG_T>(tile, in, grp0, stride);
                __syncthreads();
                ntt_round<LR1, LOG_S, COLUMN, LOG_T, false, NI>(tile, LR0, tab, log_n, g0, grp0, stride, q, ninv);
            }
        }
        ntt_round<LR0, LOG_S, COLUMN, LOG_T, false, NI, false, true, LAST && !NOSCALE, false, NOSCALE>(
            tile, 0, tab, log_n, g0, grp0, stride, q, ninv, nullptr, out);
        return;
    }

    // forward row pass: stride-1 round done in the tile, coalesced store from it
#pragma unroll
    for (int j = 0; j < G::TILE / T; ++j) {
        const int e = threadIdx.x + j * T;
        const int c = e / S, k = e % S;
        uint64_t v = tile[G::addr(c, k)];
        if constexpr (LAST) {
            v = csub_sel(v, 0 - (q << 2));
            v = csub_sel(v, 0 - (q << 1));
            v = csub_sel(v, 0 - q);
        }
        out[(grp0 * C + c) * S + k] = v;
    }
}

// XCD-aware order: workgroup b runs on XCD b % 8, so when the grid is a
// multiple of 8 each XCD gets a contiguous range of tiles (the column tiles
// that share DRAM rows, and one modulus' twiddles, stay in one L2)
__device__ __forceinline__ int xcd_block(int xcd) {
    if (xcd && (gridDim.x & 7) == 0) return (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
    return (int)blockIdx.x;
}

// strided batch: limb l of polynomial y of batch entry bi at src + bi*sB + y*sY + l*sL
template <int LOG_N, int LOG_S, int LOG_T, bool COLUMN, bool FWD, bool NOSCALE = false>
__global__ void __launch_bounds__((1 << LOG_T) >> LOG_R)
    k_ntt_pass(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
               const kbfv_tw *__restrict__ tw, const kbfv_tw *__restrict__ ninv_tab,
               const kbfv_tw *__restrict__ w1_tab, const uint64_t *__restrict__ mods, int mod0, int limbs,
               int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL, int xcd) {
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int groups = (1 << (LOG_N - LOG_S)) / G::C;  // workgroups per limb
    __shared__ uint64_t tile[G::WORDS];
    int bid = xcd_block(xcd);
    const int grp0 = bid % groups;
    bid /= groups;
    const int y = bid % polys;
    bid /= polys;
    const int bi = bid % batch;
    const int l = bid / batch;
    pass_body<LOG_N, LOG_S, LOG_T, COLUMN, FWD, NOSCALE>(tile, src + bi * sB + y * sY + l * sL,
                                                         dst + bi * dB + y * dY + l * dL, tw, ninv_tab, w1_tab,
                                                         mods, mod0 + l, grp0);
}

// ragged batch: descriptor i = {src, dst, s_limb, d_limb, mod0, limbs} (int64 words),
// limb_start[i] = the first global limb of descriptor i; the second pass (IN_DST)
// reads and writes dst
template <int LOG_N, int LOG_S, int LOG_T, bool COLUMN, bool FWD>
__global__ void __launch_bounds__((1 << LOG_T) >> LOG_R)
    k_ntt_pass_desc(const int64_t *__restrict__ desc, const int32_t *__restrict__ limb_start, int count,
                    int in_dst, const kbfv_tw *__restrict__ tw, const kbfv_tw *__restrict__ ninv_tab,
                    const uint64_t *__restrict__ mods, int xcd) {
    using G = Geo<LOG_S, COLUMN, LOG_T>;
    constexpr int groups = (1 << (LOG_N - LOG_S)) / G::C;
    __shared__ uint64_t tile[G::WORDS];
    const int bid = xcd_block(xcd);
    const int grp0 = bid % groups;
    const int gl = bid / groups;
    // the last descriptor whose first limb is at or before gl (skips empty ones)
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (limb_start[mid] <= gl)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int64_t *d = desc + 6 * (long)lo;
    const long l = gl - limb_start[lo];
    if (l >= d[5]) return;  // a total past the uploaded one: uniform per workgroup, before any barrier
    uint64_t *dst = (uint64_t *)d[1];
    const uint64_t *in = in_dst ? dst + l * d[3] : (const uint64_t *)d[0] + l * d[2];
    pass_body<LOG_N, LOG_S, LOG_T, COLUMN, FWD, false>(tile, in, dst + l * d[3], tw, ninv_tab, ninv_tab, mods,
                                                       (int)d[4] + (int)l, grp0);
}

struct pass_args {
    const kbfv_context *c;
    const uint64_t *src;
    uint64_t *dst;
    int mod0, limbs, polys, batch;
    long sB, sY, sL, dB, dY, dL;
    int xcd;
    hipStream_t st;
};

template <int LOG_N, int LOG_S, int LOG_T, bool COLUMN, bool FWD, bool NOSCALE>
int launch_pass(const pass_args &a) {
    if constexpr (!pass_ok(LOG_N, LOG_S, LOG_T, COLUMN)) {
        kb_set_error("NTT pass 2^%d over tiles of 2^%d is not in the plan space", LOG_S, LOG_T);
        return KBFV_EINVAL;
    } else {
        using G = Geo<LOG_S, COLUMN, LOG_T>;
        const long groups = (1L << (LOG_N - LOG_S)) / G::C;
        const long blocks = groups * a.limbs * a.polys * a.batch;
        KB_REQUIRE(blocks < (1L << 31), "NTT grid too large");
        hipLaunchKernelGGL((k_ntt_pass<LOG_N, LOG_S, LOG_T, COLUMN, FWD, NOSCALE>), dim3((unsigned)blocks), dim3(G::T), 0, a.st, a.src, a.dst, a.c->t.tw, a.c->t.ninv,
                           a.c->t.inv_w1, a.c->t.mod, a.mod0, a.limbs, a.polys, a.batch, a.sB, a.sY, a.sL, a.dB, a.dY,
                           a.dL, a.xcd);
        KB_LAUNCH_CHECK();
        return KBFV_OK;
    }
}

// pass length 2^log_s (one of three around the default split) and tile 2^log_t
template <int LOG_N, bool COLUMN, bool FWD, bool NOSCALE>
int launch_pass_st(int log_s, int log_t, const pass_args &a) {
    constexpr int s0 = (COLUMN ? col_log_default(LOG_N) : LOG_N - col_log_default(LOG_N)) - 1;
#define KB_NTT_ST(DS, LT)                                                   \
    if (log_s == s0 + DS && log_t == LT) return launch_pass<LOG_N, s0 + DS, LT, COLUMN, FWD, NOSCALE>(a);
    KB_NTT_ST(0, 11) KB_NTT_ST(0, 12) KB_NTT_ST(0, 13)
    KB_NTT_ST(1, 11) KB_NTT_ST(1, 12) KB_NTT_ST(1, 13)
    KB_NTT_ST(2, 11) KB_NTT_ST(2, 12) KB_NTT_ST(2, 13)
#undef KB_NTT_ST
    kb_set_error("NTT pass 2^%d over tiles of 2^%d is not in the plan space", log_s, log_t);
    return KBFV_EINVAL;
}

template <bool COLUMN, bool FWD, bool NOSCALE = false>
int launch_pass_n(int log_s, int log_t, const pass_args &a) {
    switch (a.c->log_n) {
        case 12: return launch_pass_st<12, COLUMN, FWD, NOSCALE>(log_s, log_t, a);
        case 13: return launch_pass_st<13, COLUMN, FWD, NOSCALE>(log_s, log_t, a);
        case 14: return launch_pass_st<14, COLUMN, FWD, NOSCALE>(log_s, log_t, a);
        case 15: return launch_pass_st<15, COLUMN, FWD, NOSCALE>(log_s, log_t, a);
        case 16: return launch_pass_st<16, COLUMN, FWD, NOSCALE>(log_s, log_t, a);
        default:
            kb_set_error("unsupported ring degree 2^%d", a.c->log_n);
            return KBFV_EINVAL;
    }
}

// single-pass transform of a whole limb (N <= 2^14: the limb and its row
// padding fit in the 160 KB LDS): one HBM read and one write per limb instead
// of two of each.  Register rounds: radix 16 first, the remaining log N - 4
// levels in rounds of at most 4 (12: 4,4,4; 13: 4,3,3,3; 14: 4,4,3,3).
__host__ __device__ constexpr int limb_rounds(int log_n) { return 1 + (log_n - 4 + 3) / 4; }
__host__ __device__ constexpr int limb_round_log(int log_n, int k) {
    if (k == 0) return 4;
    const int r = log_n - 4, nr = (r + 3) / 4;
    return r / nr + (k - 1 < r % nr ? 1 : 0);
}
__host__ __device__ constexpr int limb_round_lam(int log_n, int k) {
    return k == 0 ? 0 : limb_round_lam(log_n, k - 1) + limb_round_log(log_n, k - 1);
}

template <int LOG_N, int K, bool FWD, bool NI>
__device__ __forceinline__ void limb_rounds_mid(uint64_t *tile, const kbfv_tw *tab, uint64_t q, kbfv_tw ninv) {
    // rounds 1 .. limb_rounds - 1, LDS to LDS: ascending (forward) / descending (inverse)
    if constexpr (K >= 1 && K < limb_rounds(LOG_N)) {
        constexpr int KK = FWD ? K : limb_rounds(LOG_N) - K;
        ntt_round<limb_round_log(LOG_N, KK), LOG_N, false, LOG_N, FWD, NI>(tile, limb_round_lam(LOG_N, KK), tab,
                                                                         LOG_N, 0, 0, 1, q, ninv);
        limb_rounds_mid<LOG_N, K + 1, FWD, NI>(tile, tab, q, ninv);
    }
}

template <int LOG_N, bool FWD, bool NOSCALE = false>
__global__ void __launch_bounds__((1 << LOG_N) >> LOG_R)
    k_ntt_limb(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst, const kbfv_tw *__restrict__ tw,
               const kbfv_tw *__restrict__ ninv_tab, const kbfv_tw *__restrict__ w1_tab,
               const uint64_t *__restrict__ mods, int mod0, int limbs, int polys, int batch, long sB, long sY,
               long sL, long dB, long dY, long dL, int xcd) {
    static_assert(LOG_N >= 12 && LOG_N <= 14, "one limb per workgroup up to 2^14");
    using G = Geo<LOG_N, false, LOG_N>;
    constexpr int T = G::T;
    __shared__ uint64_t tile[G::WORDS];
    int bid = xcd_block(xcd);
    const int y = bid % polys;
    bid /= polys;
    const int bi = bid % batch;
    const int l = bid / batch;
    constexpr int n = 1 << LOG_N;
    const int m = mod0 + l;
    const uint64_t q = mods[m];
    const kbfv_tw *tab = tw + ((size_t)m * 2 + (FWD ? 0 : 1)) * n;
    const uint64_t *in = src + bi * sB + y * sY + l * sL;
    uint64_t *out = dst + bi * dB + y * dY + l * dL;
    constexpr bool NI = !FWD && !NOSCALE;
    const kbfv_tw ninv = NI ? ninv_tab[m] : (NOSCALE ? w1_tab[m] : kbfv_tw{0, 0});
    if constexpr (FWD) {
        // the radix-16 round of stride N/16 reads global memory in 128-byte runs
        ntt_round<4, LOG_N, false, LOG_N, true, false, true>(tile, 0, tab, LOG_N, 0, 0, 1, q, ninv, in);
        limb_rounds_mid<LOG_N, 1, true, false>(tile, tab, q, ninv);
#pragma unroll
        for (int j = 0; j < n / T; ++j) {
            const int k = threadIdx.x + j * T;
            uint64_t v = tile[G::addr(0, k)];
            v = csub_sel(v, 0 - (q << 2));
            v = csub_sel(v, 0 - (q << 1));
            v = csub_sel(v, 0 - q);
            out[k] = v;
        }
    } else {
        stage_in<LOG_N, false, LOG_N>(tile, in, 0, 1);
        __syncthreads();
        limb_rounds_mid<LOG_N, 1, false, NI>(tile, tab, q, ninv);
        ntt_round<4, LOG_N, false, LOG_N, false, NI, false, true, !NOSCALE, false, NOSCALE>(
            tile, 0, tab, LOG_N, 0, 0, 1, q, ninv, nullptr, out);
    }
}

// forward column pass of the external product's digit polynomials, the digit
// split fused into the load: row r = (h, i, j) of the gadget takes digit j of
// limb i of component h of the coefficient-form ciphertext, reduced to every
// Q limb l.  coef [batch][2][qn][N] -> dig [batch][rc][qn][N] (rows r0..r0+rc-1);
// the target limb varies fastest, so the qn workgroups that read one source
// tile run back to back on one XCD and share it through L2
template <int LOG_N, int LOG_S, int LOG_T>
__global__ void __launch_bounds__((1 << LOG_T) >> LOG_R)
    k_ntt_digit_col(const uint64_t *__restrict__ coef, uint64_t *__restrict__ dig, const kbfv_tw *__restrict__ tw,
                    const uint64_t *__restrict__ mods, int oQ, int qn, int nd, int bits, int r0, int rc, int xcd) {
    using G = Geo<LOG_S, true, LOG_T>;
    constexpr int groups = (1 << (LOG_N - LOG_S)) / G::C;
    constexpr long n = 1L << LOG_N;
    __shared__ uint64_t tile[G::WORDS];
    int bid = xcd_block(xcd);
    const int l = bid % qn;
    bid /= qn;
    const int grp0 = bid % groups;
    bid /= groups;
    const int y = bid % rc;
    const long bt = bid / rc;
    const int row = r0 + y;
    const long L = qn * n;
    const uint64_t *in = coef + bt * 2 * L + (long)(row / nd) * n;  // (h*qn + i)*N = h*L + i*N
    uint64_t *out = dig + (bt * rc + y) * L + (long)l * n;
    pass_body<LOG_N, LOG_S, LOG_T, true, true, false, true>(tile, in, out, tw, nullptr, nullptr, mods, oQ + l, grp0,
                                                            bits * (row % nd), (1ULL << bits) - 1);
}

// ---------------------------------------------------------------- plans
struct plan_fields {
    int col_s, col_t, row_t, col_x, row_x;
    int chunk;  // 0: each pass over the whole launch; k: chunks of at most 2^(k+23) bytes
};

plan_fields decode_plan(int log_n, uint32_t p) {
    if (p == 0) return {col_log_default(log_n), 12, 12, 1, 1, 0};
    return {(int)(p & 15),         (int)((p >> 4) & 15), (int)((p >> 8) & 15),
            (int)((p >> 12) & 1),  (int)((p >> 13) & 1), (int)((p >> 14) & 7)};
}

// single-pass plan: column length = log N, no tiles, row order bit clear
bool single_pass(int log_n, const plan_fields &f) { return f.col_s == log_n; }

bool plan_ok(int log_n, uint32_t p) {
    if (p == 0) return true;
    if (p >> 17) return false;
    const plan_fields f = decode_plan(log_n, p);
    if (single_pass(log_n, f))
        return log_n <= 14 && f.col_t == 0 && f.row_t == 0 && f.row_x == 0 && f.chunk == 0;
    if (f.chunk != 0 && (f.chunk < 2 || f.chunk > 6)) return false;
    const int d = f.col_s - col_log_default(log_n);
    return d >= -1 && d <= 1 && pass_ok(log_n, f.col_s, f.col_t, true) &&
           pass_ok(log_n, log_n - f.col_s, f.row_t, false);
}

uint32_t encode_plan(int col_s, int col_t, int row_t, int col_x, int row_x, int chunk = 0) {
    return (uint32_t)col_s | (uint32_t)col_t << 4 | (uint32_t)row_t << 8 | (uint32_t)col_x << 12 |
           (uint32_t)row_x << 13 | (uint32_t)chunk << 14;
}

int bucket_of(long transforms) {
    return transforms < 64 ? 0 : transforms < 512 ? 1 : transforms < 4096 ? 2 : 3;
}

uint32_t plan_for(const kbfv_context *c, int dir, long transforms) {
    return kb_root(c)->ntt_plan[dir > 0 ? 0 : 1][bucket_of(transforms)].load(std::memory_order_relaxed);
}

int check_args(const kbfv_context *c, const uint64_t *src, const uint64_t *dst, int mod0, int limbs, int polys,
               int batch) {
    KB_REQUIRE(c && src && dst, "null argument");
    KB_REQUIRE(c->log_n >= 12 && c->log_n <= 16, "NTT supports 2^12..2^16");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod, "modulus range out of bounds");
    KB_REQUIRE(polys > 0 && batch > 0, "empty NTT");
    return KBFV_OK;
}

// one transform with an explicit plan; NOSCALE: inverse without n^-1 and the final reduction
template <int LOG_N, bool FWD, bool NOSCALE>
int launch_limb(const pass_args &a) {
    const long blocks = (long)a.limbs * a.polys * a.batch;
    KB_REQUIRE(blocks < (1L << 31), "NTT grid too large");
    hipLaunchKernelGGL((k_ntt_limb<LOG_N, FWD, NOSCALE>), dim3((unsigned)blocks), dim3((1 << LOG_N) >> LOG_R), 0, a.st,
                       a.src, a.dst, a.c->t.tw, a.c->t.ninv, a.c->t.inv_w1, a.c->t.mod, a.mod0, a.limbs, a.polys,
                       a.batch, a.sB, a.sY, a.sL, a.dB, a.dY, a.dL, a.xcd);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

template <bool FWD, bool NOSCALE>
int launch_limb_n(const pass_args &a) {
    switch (a.c->log_n) {
        case 12: return launch_limb<12, FWD, NOSCALE>(a);
        case 13: return launch_limb<13, FWD, NOSCALE>(a);
        case 14: return launch_limb<14, FWD, NOSCALE>(a);
        default:
            kb_set_error("single-pass NTT needs N <= 2^14");
            return KBFV_EINVAL;
    }
}

int ntt_with_plan(const kbfv_context *c, int dir, bool noscale, uint32_t plan, const uint64_t *src, uint64_t *dst,
                  int mod0, int limbs, int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL,
                  hipStream_t st) {
    const plan_fields f = decode_plan(c->log_n, plan);
    if (single_pass(c->log_n, f)) {
        const pass_args a{c, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, f.col_x, st};
        if (dir > 0) return launch_limb_n<true, false>(a);
        return noscale ? launch_limb_n<false, true>(a) : launch_limb_n<false, false>(a);
    }
    const int row_s = c->log_n - f.col_s;
    // both passes of one chunk of the launch before the next chunk: a chunk
    // small enough for the 256 MB last-level (Infinity) cache is still there
    // when its second pass reads what its first pass wrote.  The chunk runs
    // over the outermost dimension with more than one entry (batch, then
    // polynomials, then limbs).
    const long limb_bytes = (long)c->n * 8;
    const long cap = f.chunk ? (1L << (f.chunk + 23)) : 0;
    int dim = 0, total = 1, step = 1;
    if (cap) {
        if (batch > 1)
            dim = 0, total = batch, step = (int)std::max(1L, cap / (limb_bytes * limbs * polys));
        else if (polys > 1)
            dim = 1, total = polys, step = (int)std::max(1L, cap / (limb_bytes * limbs));
        else
            dim = 2, total = limbs, step = (int)std::max(1L, cap / limb_bytes);
        if (step >= total) step = total;
    }
    for (int k0 = 0; k0 < total; k0 += step) {
        const int kn = std::min(step, total - k0);
        int cb = batch, cp = polys, cl = limbs, cm = mod0;
        long so = 0, dof = 0;
        if (cap && dim == 0) cb = kn, so = k0 * sB, dof = k0 * dB;
        if (cap && dim == 1) cp = kn, so = k0 * sY, dof = k0 * dY;
        if (cap && dim == 2) cl = kn, cm = mod0 + k0, so = k0 * sL, dof = k0 * dL;
        pass_args first{c, src + so, dst + dof, cm, cl, cp, cb, sB, sY, sL, dB, dY, dL,
                        dir > 0 ? f.col_x : f.row_x, st};
        pass_args second{c, dst + dof, dst + dof, cm, cl, cp, cb, dB, dY, dL, dB, dY, dL,
                         dir > 0 ? f.row_x : f.col_x, st};
        if (dir > 0) {
            KB_TRY((launch_pass_n<true, true>(f.col_s, f.col_t, first)));
            KB_TRY((launch_pass_n<false, true>(row_s, f.row_t, second)));
            continue;
        }
        KB_TRY((launch_pass_n<false, false>(row_s, f.row_t, first)));
        if (noscale)
            KB_TRY((launch_pass_n<true, false, true>(f.col_s, f.col_t, second)));
        else
            KB_TRY((launch_pass_n<true, false>(f.col_s, f.col_t, second)));
    }
    return KBFV_OK;
}

}  // namespace

int kb_ntt(const kbfv_context *c, int dir, const uint64_t *src, uint64_t *dst, int mod0, int limbs,
           int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL,
           hipStream_t st) {
    KB_TRY(check_args(c, src, dst, mod0, limbs, polys, batch));
    if (mod0 + limbs - 1 == c->oPlain) KB_REQUIRE(c->batching, "plain modulus has no 2N-th root");
    const uint32_t plan = plan_for(c, dir, (long)limbs * polys * batch);
    return ntt_with_plan(c, dir, false, plan, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, st);
}

int kb_ntt_inplace(const kbfv_context *c, int dir, uint64_t *data, int mod0, int limbs, int polys,
                   hipStream_t st) {
    const long N = c->n;
    return kb_ntt(c, dir, data, data, mod0, limbs, polys, 1, 0, (long)limbs * N, N, 0,
                  (long)limbs * N, N, st);
}

int kb_intt_noscale(const kbfv_context *c, const uint64_t *src, uint64_t *dst, int mod0, int limbs, int polys,
                    int batch, long sB, long sY, long sL, long dB, long dY, long dL, hipStream_t st) {
    KB_TRY(check_args(c, src, dst, mod0, limbs, polys, batch));
    const uint32_t plan = plan_for(c, -1, (long)limbs * polys * batch);
    return ntt_with_plan(c, -1, true, plan, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, st);
}

namespace {
template <int LOG_N>
int launch_digit_col(const kbfv_context *c, int log_t, const uint64_t *coef, uint64_t *dig, int batch, int bits, int nd,
                     int r0, int rc, hipStream_t st) {
    constexpr int LOG_S = col_log_default(LOG_N);
    auto go = [&](auto tag) -> int {
        constexpr int LT = decltype(tag)::value;
        if constexpr (!pass_ok(LOG_N, LOG_S, LT, true)) {
            kb_set_error("digit NTT: tile 2^%d does not fit N=2^%d", LT, LOG_N);
            return KBFV_EINVAL;
        } else {
            using G = Geo<LOG_S, true, LT>;
            const long blocks = (long)c->qn * ((1L << (LOG_N - LOG_S)) / G::C) * rc * batch;
            KB_REQUIRE(blocks < (1L << 31), "NTT grid too large");
            hipLaunchKernelGGL((k_ntt_digit_col<LOG_N, LOG_S, LT>), dim3((unsigned)blocks), dim3(G::T), 0, st, coef,
                               dig, c->t.tw, c->t.mod, c->oQ, c->qn, nd, bits, r0, rc, 1);
            KB_LAUNCH_CHECK();
            return KBFV_OK;
        }
    };
    switch (log_t) {
        case 11: return go(std::integral_constant<int, 11>{});
        case 12: return go(std::integral_constant<int, 12>{});
        default: return go(std::integral_constant<int, 13>{});
    }
}
}  // namespace

int kb_digit_ntt(const kbfv_context *c, const uint64_t *coef, uint64_t *dig, int batch, int bits, int nd, int r0,
                 int rc, hipStream_t st) {
    KB_REQUIRE(c && coef && dig && batch > 0 && rc > 0 && nd > 0 && bits >= 1 && bits <= 62, "bad digit NTT arguments");
    KB_REQUIRE(c->log_n >= 12 && c->log_n <= 16, "NTT supports 2^12..2^16");
    const int qn = c->qn, log_n = c->log_n;
    const long N = c->n, L = (long)qn * N;
    // the fused split runs the built-in split; it keeps the plan's tiles and
    // row order when the plan uses that split too
    const plan_fields f = decode_plan(log_n, plan_for(c, +1, (long)qn * rc * batch));
    const bool same = f.col_s == col_log_default(log_n);
    const int col_t = same ? f.col_t : 12, row_t = same ? f.row_t : 12, row_x = same ? f.row_x : 1;
    int rc2;
    switch (log_n) {
        case 12: rc2 = launch_digit_col<12>(c, col_t, coef, dig, batch, bits, nd, r0, rc, st); break;
        case 13: rc2 = launch_digit_col<13>(c, col_t, coef, dig, batch, bits, nd, r0, rc, st); break;
        case 14: rc2 = launch_digit_col<14>(c, col_t, coef, dig, batch, bits, nd, r0, rc, st); break;
        case 15: rc2 = launch_digit_col<15>(c, col_t, coef, dig, batch, bits, nd, r0, rc, st); break;
        default: rc2 = launch_digit_col<16>(c, col_t, coef, dig, batch, bits, nd, r0, rc, st); break;
    }
    if (rc2) return rc2;
    const pass_args second{c, dig, dig, c->oQ, qn, rc, batch, rc * L, L, N, rc * L, L, N, row_x, st};
    return launch_pass_n<false, true>(log_n - col_log_default(log_n), row_t, second);
}

// ---------------------------------------------------------------- plan ABI and autotuner
extern "C" int kbfv_ntt_plan_valid(int log_n, uint32_t plan) {
    KB_REQUIRE(log_n >= 12 && log_n <= 16, "NTT supports 2^12..2^16");
    return plan_ok(log_n, plan) ? KBFV_OK : KBFV_EINVAL;
}

extern "C" int kbfv_ntt_plan_candidates(int log_n, uint32_t *out, int cap) {
    KB_REQUIRE(log_n >= 12 && log_n <= 16, "NTT supports 2^12..2^16");
    KB_REQUIRE(cap >= 0 && (out || cap == 0), "null output");
    int k = 0;
    for (int cx = 0; cx <= 1 && log_n <= 14; ++cx) {
        const uint32_t p = encode_plan(log_n, 0, 0, cx, 0);
        if (k < cap) out[k] = p;
        ++k;
    }
    for (int d = -1; d <= 1; ++d)
        for (int ct = 11; ct <= 13; ++ct)
            for (int rt = 11; rt <= 13; ++rt)
                for (int cx = 0; cx <= 1; ++cx)
                    for (int rx = 0; rx <= 1; ++rx) {
                        const uint32_t p = encode_plan(col_log_default(log_n) + d, ct, rt, cx, rx);
                        if (!plan_ok(log_n, p)) continue;
                        if (k < cap) out[k] = p;
                        ++k;
                    }
    return k;
}

extern "C" int kbfv_ntt_plan_bucket(int64_t transforms) {
    KB_REQUIRE(transforms > 0, "empty NTT");
    return bucket_of((long)transforms);
}

extern "C" int kbfv_ntt_plan_get(const kbfv_context *c, int dir, int bucket, uint32_t *plan) {
    KB_REQUIRE(c && plan, "null argument");
    KB_REQUIRE(dir == 1 || dir == -1, "dir must be +1 or -1");
    KB_REQUIRE(bucket >= 0 && bucket < KBFV_NTT_BUCKETS, "bucket out of range");
    *plan = kb_root(c)->ntt_plan[dir > 0 ? 0 : 1][bucket].load(std::memory_order_relaxed);
    return KBFV_OK;
}

extern "C" int kbfv_ntt_plan_set(kbfv_context *c, int dir, int bucket, uint32_t plan) {
    KB_REQUIRE(c, "null argument");
    KB_REQUIRE(dir == 1 || dir == -1, "dir must be +1 or -1");
    KB_REQUIRE(bucket >= 0 && bucket < KBFV_NTT_BUCKETS, "bucket out of range");
    KB_REQUIRE(plan_ok(c->log_n, plan), "NTT plan 0x%x is not valid at N=2^%d", plan, c->log_n);
    // a level view shares the plans of its full-level context
    const_cast<kbfv_context *>(kb_root(c))->ntt_plan[dir > 0 ? 0 : 1][bucket].store(plan, std::memory_order_relaxed);
    return KBFV_OK;
}

namespace {
// average milliseconds of one pass variant over `reps` launches after one warm-up
int time_pass(const kbfv_context *c, int dir, bool column, int log_s, int log_t, int xcd, uint64_t *buf, int mod0,
              int limbs, int polys, int reps, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float *ms) {
    const long N = c->n;
    const pass_args a{c, buf, buf, mod0, limbs, polys, 1, 0, (long)limbs * N, N, 0, (long)limbs * N, N, xcd, st};
    auto run = [&]() -> int {
        if (column) return dir > 0 ? launch_pass_n<true, true>(log_s, log_t, a) : launch_pass_n<true, false>(log_s, log_t, a);
        return dir > 0 ? launch_pass_n<false, true>(log_s, log_t, a) : launch_pass_n<false, false>(log_s, log_t, a);
    };
    KB_TRY(run());
    KB_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) KB_TRY(run());
    KB_HIP(hipEventRecord(e1, st));
    KB_HIP(hipEventSynchronize(e1));
    float t = 0;
    KB_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return KBFV_OK;
}
}  // namespace

extern "C" int kbfv_ntt_autotune(kbfv_context *c, int dir, int mod0, int limbs, int polys, uint64_t *scratch,
                                 int reps, void *stream, uint32_t *best, float *best_ms) {
    KB_REQUIRE(dir == 1 || dir == -1, "dir must be +1 or -1");
    KB_TRY(check_args(c, scratch, scratch, mod0, limbs, polys, 1));
    KB_REQUIRE(mod0 + limbs - 1 != c->oPlain, "tune over RNS limbs, not the plain modulus");
    KB_REQUIRE(reps >= 1 && reps <= 100, "reps must be in [1, 100]");
    hipStream_t st = (hipStream_t)stream;
    const int log_n = c->log_n;
    // residues of 0: the passes read and write every element whatever its value
    KB_HIP(hipMemsetAsync(scratch, 0, (size_t)polys * limbs * c->n * sizeof(uint64_t), st));
    hipEvent_t e0, e1;
    KB_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        kb_set_error("hipEventCreate failed");
        return KBFV_EHIP;
    }
    int rc = KBFV_OK;
    float best_total = 1e30f;
    uint32_t best_plan = 0;
    // the two passes are separate launches over the same data: tune each pass's
    // tile and grid order for every split, then take the split with the least sum
    for (int d = -1; d <= 1 && rc == KBFV_OK; ++d) {
        const int cs = col_log_default(log_n) + d, rs = log_n - cs;
        float bc = 1e30f, br = 1e30f;
        int bct = 0, bcx = 0, brt = 0, brx = 0;
        for (int t = 11; t <= 13 && rc == KBFV_OK; ++t)
            for (int x = 0; x <= 1 && rc == KBFV_OK; ++x) {
                float ms;
                if (pass_ok(log_n, cs, t, true)) {
                    rc = time_pass(c, dir, true, cs, t, x, scratch, mod0, limbs, polys, reps, st, e0, e1, &ms);
                    if (rc == KBFV_OK && ms < bc) bc = ms, bct = t, bcx = x;
                }
                if (rc == KBFV_OK && pass_ok(log_n, rs, t, false)) {
                    rc = time_pass(c, dir, false, rs, t, x, scratch, mod0, limbs, polys, reps, st, e0, e1, &ms);
                    if (rc == KBFV_OK && ms < br) br = ms, brt = t, brx = x;
                }
            }
        if (rc == KBFV_OK && bct && brt && bc + br < best_total) {
            best_total = bc + br;
            best_plan = encode_plan(cs, bct, brt, bcx, brx);
        }
    }
    // one launch per transform where the limb fits a workgroup's LDS
    for (int x = 0; x <= 1 && rc == KBFV_OK && log_n <= 14; ++x) {
        const long N = c->n;
        const pass_args a{c, scratch, scratch, mod0, limbs, polys, 1, 0, (long)limbs * N, N, 0, (long)limbs * N, N, x,
                          st};
        auto run = [&]() -> int { return dir > 0 ? launch_limb_n<true, false>(a) : launch_limb_n<false, false>(a); };
        rc = run();
        if (rc == KBFV_OK) rc = hipEventRecord(e0, st) == hipSuccess ? KBFV_OK : KBFV_EHIP;
        for (int r = 0; r < reps && rc == KBFV_OK; ++r) rc = run();
        if (rc == KBFV_OK) rc = hipEventRecord(e1, st) == hipSuccess ? KBFV_OK : KBFV_EHIP;
        if (rc == KBFV_OK) rc = hipEventSynchronize(e1) == hipSuccess ? KBFV_OK : KBFV_EHIP;
        float t = 0;
        if (rc == KBFV_OK) rc = hipEventElapsedTime(&t, e0, e1) == hipSuccess ? KBFV_OK : KBFV_EHIP;
        if (rc == KBFV_EHIP) kb_set_error("NTT autotune: HIP event call failed");
        if (rc == KBFV_OK && t / reps < best_total) {
            best_total = t / reps;
            best_plan = encode_plan(log_n, 0, 0, x, 0);
        }
    }
    // chunked two-pass transforms (both passes per chunk of at most 2^(k+23)
    // bytes) with the picked tiles, timed as whole transforms against the
    // unchunked one
    if (rc == KBFV_OK && best_plan != 0 && !single_pass(log_n, decode_plan(log_n, best_plan))) {
        const long N = c->n;
        const long bytes = (long)polys * limbs * N * 8;
        auto time_transform = [&](uint32_t plan, float *ms) -> int {
            auto run = [&]() -> int {
                return ntt_with_plan(c, dir, false, plan, scratch, scratch, mod0, limbs, polys, 1, 0,
                                     (long)limbs * N, N, 0, (long)limbs * N, N, st);
            };
            KB_TRY(run());
            KB_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r) KB_TRY(run());
            KB_HIP(hipEventRecord(e1, st));
            KB_HIP(hipEventSynchronize(e1));
            float t = 0;
            KB_HIP(hipEventElapsedTime(&t, e0, e1));
            *ms = t / reps;
            return KBFV_OK;
        };
        float whole = 0;
        rc = time_transform(best_plan, &whole);
        if (rc == KBFV_OK) best_total = whole;
        for (int k = 2; k <= 6 && rc == KBFV_OK; ++k) {
            if ((1L << (k + 23)) >= bytes) break;  // one chunk: the same launches as unchunked
            float ms = 0;
            const uint32_t p = (best_plan & 0x3fffu) | (uint32_t)k << 14;
            rc = time_transform(p, &ms);
            if (rc == KBFV_OK && ms < best_total) best_total = ms, best_plan = p;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != KBFV_OK) return rc;
    KB_REQUIRE(best_plan != 0, "no NTT plan fits N=2^%d", log_n);
    const_cast<kbfv_context *>(kb_root(c))->ntt_plan[dir > 0 ? 0 : 1][bucket_of((long)limbs * polys)].store(
        best_plan, std::memory_order_relaxed);
    if (best) *best = best_plan;
    if (best_ms) *best_ms = best_total;
    return KBFV_OK;
}

// ---------------------------------------------------------------- ragged batches
namespace {
template <int LOG_N, bool COLUMN, bool FWD>
int launch_desc(const kbfv_context *c, const int64_t *desc, const int32_t *start, int count, int total, int in_dst,
                hipStream_t st) {
    constexpr int LOG_S = COLUMN ? col_log_default(LOG_N) : LOG_N - col_log_default(LOG_N);
    using G = Geo<LOG_S, COLUMN, 12>;
    const long blocks = (long)total * ((1L << (LOG_N - LOG_S)) / G::C);
    KB_REQUIRE(blocks < (1L << 31), "NTT grid too large");
    hipLaunchKernelGGL((k_ntt_pass_desc<LOG_N, LOG_S, 12, COLUMN, FWD>), dim3((unsigned)blocks), dim3(G::T), 0, st,
                       desc, start, count, in_dst, c->t.tw, c->t.ninv, c->t.mod, 1);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

template <bool COLUMN, bool FWD>
int launch_desc_n(const kbfv_context *c, const int64_t *desc, const int32_t *start, int count, int total,
                  int in_dst, hipStream_t st) {
    switch (c->log_n) {
        case 12: return launch_desc<12, COLUMN, FWD>(c, desc, start, count, total, in_dst, st);
        case 13: return launch_desc<13, COLUMN, FWD>(c, desc, start, count, total, in_dst, st);
        case 14: return launch_desc<14, COLUMN, FWD>(c, desc, start, count, total, in_dst, st);
        case 15: return launch_desc<15, COLUMN, FWD>(c, desc, start, count, total, in_dst, st);
        case 16: return launch_desc<16, COLUMN, FWD>(c, desc, start, count, total, in_dst, st);
        default:
            kb_set_error("unsupported ring degree 2^%d", c->log_n);
            return KBFV_EINVAL;
    }
}
}  // namespace

extern "C" int kbfv_ntt_desc_upload(const kbfv_context *c, const int64_t *host_desc, int count, int64_t *dev_desc,
                                    int32_t *dev_limb_start, int *total_limbs, void *stream) {
    KB_REQUIRE(c && host_desc && dev_desc && dev_limb_start && total_limbs, "null argument");
    KB_REQUIRE(count >= 1 && count <= (1 << 20), "descriptor count out of range");
    std::vector<int32_t> start((size_t)count + 1);
    long total = 0;
    for (int i = 0; i < count; ++i) {
        const int64_t *d = host_desc + 6 * (size_t)i;
        const int64_t mod0 = d[4], limbs = d[5];
        KB_REQUIRE(limbs >= 0 && mod0 >= 0 && mod0 + limbs <= c->nmod, "descriptor %d: modulus range out of bounds", i);
        KB_REQUIRE(limbs == 0 || (d[0] && d[1]), "descriptor %d: null polynomial", i);
        KB_REQUIRE(limbs == 0 || mod0 + limbs - 1 != c->oPlain || c->batching,
                   "descriptor %d: plain modulus has no 2N-th root", i);
        KB_REQUIRE(limbs <= 1 || (d[2] >= c->n && d[3] >= c->n), "descriptor %d: limbs overlap", i);
        start[i] = (int32_t)total;
        total += limbs;
        KB_REQUIRE(total * c->n < (1L << 40), "ragged batch too large");
    }
    KB_REQUIRE(total > 0, "empty NTT");
    start[count] = (int32_t)total;
    hipStream_t st = (hipStream_t)stream;
    KB_HIP(hipMemcpyAsync(dev_desc, host_desc, (size_t)count * 6 * sizeof(int64_t), hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(dev_limb_start, start.data(), start.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    // the host vector goes out of scope: wait for the copies to leave it
    KB_HIP(hipStreamSynchronize(st));
    *total_limbs = (int)total;
    return KBFV_OK;
}

extern "C" int kbfv_ntt_desc(const kbfv_context *c, int dir, const int64_t *dev_desc, const int32_t *dev_limb_start,
                             int count, int total_limbs, void *stream) {
    KB_REQUIRE(c && dev_desc && dev_limb_start, "null argument");
    KB_REQUIRE(dir == 1 || dir == -1, "dir must be +1 or -1");
    KB_REQUIRE(count >= 1 && total_limbs >= 1, "empty NTT");
    KB_REQUIRE(c->log_n >= 12 && c->log_n <= 16, "NTT supports 2^12..2^16");
    hipStream_t st = (hipStream_t)stream;
    if (dir > 0) {
        KB_TRY((launch_desc_n<true, true>(c, dev_desc, dev_limb_start, count, total_limbs, 0, st)));
        return launch_desc_n<false, true>(c, dev_desc, dev_limb_start, count, total_limbs, 1, st);
    }
    KB_TRY((launch_desc_n<false, false>(c, dev_desc, dev_limb_start, count, total_limbs, 0, st)));
    return launch_desc_n<true, false>(c, dev_desc, dev_limb_start, count, total_limbs, 1, st);
}

// ---------------------------------------------------------------- ragged element-wise
namespace {
constexpr int RB_BS = 256;

// descriptor i = {a, b, out, s_a, s_b, s_out, mod0, limbs} (int64 words); op 0
// add, 1 sub, 2 mul, on canonical residues; workgroup = 256 coefficients of one limb
__global__ void __launch_bounds__(RB_BS)
    k_binop_desc(const int64_t *__restrict__ desc, const int32_t *__restrict__ limb_start, int count, int op,
                 const uint64_t *__restrict__ mods, const uint64_t *__restrict__ mu, int chunks) {
    const int gl = (int)(blockIdx.x / chunks);
    const int k = (int)(blockIdx.x % chunks) * RB_BS + threadIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (limb_start[mid] <= gl)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int64_t *d = desc + 8 * (long)lo;
    const long l = gl - limb_start[lo];
    if (l >= d[7]) return;  // a total past the uploaded one
    const int m = (int)d[6] + (int)l;
    const uint64_t q = mods[m];
    const uint64_t a = ((const uint64_t *)d[0])[l * d[3] + k];
    const uint64_t b = ((const uint64_t *)d[1])[l * d[4] + k];
    uint64_t r;
    if (op == 0)
        r = kb_addmod(a, b, q);
    else if (op == 1)
        r = kb_submod(a, b, q);
    else
        r = kb_mulmod(a, b, q, mu[2 * m], mu[2 * m + 1]);
    ((uint64_t *)d[2])[l * d[5] + k] = r;
}
}  // namespace

extern "C" int kbfv_poly_binop_desc_upload(const kbfv_context *c, const int64_t *host_desc, int count,
                                           int64_t *dev_desc, int32_t *dev_limb_start, int *total_limbs,
                                           void *stream) {
    KB_REQUIRE(c && host_desc && dev_desc && dev_limb_start && total_limbs, "null argument");
    KB_REQUIRE(count >= 1 && count <= (1 << 20), "descriptor count out of range");
    std::vector<int32_t> start((size_t)count + 1);
    long total = 0;
    for (int i = 0; i < count; ++i) {
        const int64_t *d = host_desc + 8 * (size_t)i;
        const int64_t mod0 = d[6], limbs = d[7];
        KB_REQUIRE(limbs >= 0 && mod0 >= 0 && mod0 + limbs <= c->nmod, "descriptor %d: modulus range out of bounds", i);
        KB_REQUIRE(limbs == 0 || (d[0] && d[1] && d[2]), "descriptor %d: null polynomial", i);
        KB_REQUIRE(limbs <= 1 || (d[3] >= c->n && d[4] >= c->n && d[5] >= c->n), "descriptor %d: limbs overlap", i);
        start[i] = (int32_t)total;
        total += limbs;
        KB_REQUIRE(total * c->n < (1L << 40), "ragged batch too large");
    }
    KB_REQUIRE(total > 0, "empty batch");
    start[count] = (int32_t)total;
    hipStream_t st = (hipStream_t)stream;
    KB_HIP(hipMemcpyAsync(dev_desc, host_desc, (size_t)count * 8 * sizeof(int64_t), hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(dev_limb_start, start.data(), start.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    KB_HIP(hipStreamSynchronize(st));  // the host vector goes out of scope
    *total_limbs = (int)total;
    return KBFV_OK;
}

extern "C" int kbfv_poly_binop_desc(const kbfv_context *c, int op, const int64_t *dev_desc,
                                    const int32_t *dev_limb_start, int count, int total_limbs, void *stream) {
    KB_REQUIRE(c && dev_desc && dev_limb_start, "null argument");
    KB_REQUIRE(op >= 0 && op <= 2, "op must be 0 (add), 1 (sub) or 2 (mul)");
    KB_REQUIRE(count >= 1 && total_limbs >= 1, "empty batch");
    const int chunks = c->n / RB_BS;
    const long blocks = (long)total_limbs * chunks;
    KB_REQUIRE(blocks < (1L << 31), "grid too large");
    hipLaunchKernelGGL(k_binop_desc, dim3((unsigned)blocks), dim3(RB_BS), 0, (hipStream_t)stream, dev_desc,
                       dev_limb_start, count, op, c->t.mod, c->t.mu, chunks);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
