// Batched negacyclic NTT / INTT over 60-bit primes for gfx950.
//
// Computes the same transform as the reference's two-kernel radix-8 NTT
// (include/ntt_60bit.cuh:414-823, dispatch include/BFVcontext.cuh:1513-1552):
// Cooley-Tukey forward with the bit-reversed psi table (stage with m groups uses
// psi[m + group]), Gentleman-Sande inverse with the bit-reversed psi^-1 table
// whose entry 1 carries n^-1, outputs canonical in [0, q).
//
// Design (MI355X): N = N1 * N2 four-step split into two passes, each a kernel
// that stages a 32 KiB tile in LDS and runs radix-16 rounds in registers:
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
#include "kbfv_internal.h"

namespace {

constexpr int LOG_R = 4;  // radix-16 register rounds
constexpr int TILE = 4096;  // elements per workgroup

template <int LOG_S, bool COLUMN>
struct Geo {
    static constexpr int S = 1 << LOG_S;
    static constexpr int C = TILE / S;         // sub-transforms per workgroup
    static constexpr int T = TILE >> LOG_R;    // threads per workgroup (256)
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
template <int LR, int LOG_S, bool COLUMN>
__device__ __forceinline__ void round_twiddles(kbfv_tw *tws, const int lam0, const kbfv_tw *tab,
                                               const int log_n, const int g0, const int grp0,
                                               const int stride) {
    using G = Geo<LOG_S, COLUMN>;
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
template <int LR, int LOG_S, bool COLUMN, bool FWD, bool NINV, bool GIN = false, bool GOUT = false,
          bool FIN = false, bool PRE = false, bool W1 = false>
__device__ __forceinline__ void ntt_round(uint64_t *tile, const int lam0, const kbfv_tw *tab,
                                          const int log_n, const int g0, const int grp0,
                                          const int stride, const uint64_t q, const kbfv_tw ninv,
                                          const uint64_t *__restrict__ gin = nullptr,
                                          uint64_t *__restrict__ gout = nullptr,
                                          const kbfv_tw (*pre)[(1 << LR) - 1] = nullptr) {
    using G = Geo<LOG_S, COLUMN>;
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
            if constexpr (GIN)
                x[i] = gin[posc + (kbase + i * d) * pstep];
            else
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
template <int LOG_S, bool COLUMN>
__device__ __forceinline__ void stage_in(uint64_t *tile, const uint64_t *__restrict__ in, int grp0,
                                         int stride) {
    using G = Geo<LOG_S, COLUMN>;
    constexpr int S = G::S, C = G::C, T = G::T;
#pragma unroll
    for (int j = 0; j < TILE / T; ++j) {
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

// pass sizes: the column pass takes ceil(log_n / 2) levels, the row pass the rest
template <int LOG_N, bool COLUMN>
constexpr int pass_log() { return COLUMN ? (LOG_N + 1) / 2 : LOG_N - (LOG_N + 1) / 2; }

template <int LOG_N, bool COLUMN, bool FWD, bool NOSCALE = false>
__global__ void __launch_bounds__(TILE >> LOG_R)
    k_ntt_pass(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
               const kbfv_tw *__restrict__ tw, const kbfv_tw *__restrict__ ninv_tab,
               const kbfv_tw *__restrict__ w1_tab, const uint64_t *__restrict__ mods, int mod0, int limbs,
               int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL) {
    constexpr int LOG_S = pass_log<LOG_N, COLUMN>();
    constexpr int log_n = LOG_N;
    using G = Geo<LOG_S, COLUMN>;
    constexpr int S = G::S, C = G::C, T = G::T;
    constexpr int groups = (1 << (LOG_N - LOG_S)) / C;  // workgroups per limb
    static_assert(groups >= 1, "ring too small for the tile");
    constexpr bool LAST = (FWD != COLUMN);  // forward ends with the row pass
    __shared__ uint64_t tile[G::WORDS];

    // XCD-aware order: workgroup b runs on XCD b % 8, so when the grid is a
    // multiple of 8 each XCD gets a contiguous range of tiles (the column
    // tiles that share DRAM rows, and one modulus' twiddles, stay in one L2)
    int bid = (int)blockIdx.x;
    if ((gridDim.x & 7) == 0) bid = (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
    const int grp0 = bid % groups;
    bid /= groups;
    const int y = bid % polys;
    bid /= polys;
    const int bi = bid % batch;
    const int l = bid / batch;

    constexpr int n = 1 << log_n;
    const int m = mod0 + l;
    const uint64_t q = mods[m];
    const kbfv_tw *tab = tw + ((size_t)m * 2 + (FWD ? 0 : 1)) * n;
    const uint64_t *in = src + bi * sB + y * sY + l * sL;
    uint64_t *out = dst + bi * dB + y * dY + l * dL;
    constexpr int stride = n >> LOG_S;
    constexpr int g0 = COLUMN ? 0 : (log_n - LOG_S);

    constexpr int LR0 = LOG_S < LOG_R ? LOG_S : LOG_R;
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
        ntt_round<LR0, LOG_S, COLUMN, true, false, true>(tile, 0, tab, log_n, g0, grp0, stride, q, ninv, in);
        if constexpr (COLUMN) {
            ntt_round<LR1, LOG_S, COLUMN, true, false, false, true, LAST>(tile, LR0, tab, log_n, g0, grp0,
                                                                          stride, q, ninv, nullptr, out);
            return;
        } else {
            ntt_round<LR1, LOG_S, COLUMN, true, false>(tile, LR0, tab, log_n, g0, grp0, stride, q, ninv);
        }
    } else {
        if constexpr (COLUMN) {
            ntt_round<LR1, LOG_S, COLUMN, false, NI, true>(tile, LR0, tab, log_n, g0, grp0, stride, q, ninv,
                                                           in);
        } else {
            // one butterfly group per thread (a radix-16 round): its twiddles are
            // loaded with the tile, ahead of the barrier, rather than after it
            constexpr bool PRE = (C * (S >> LR1) == T);
            if constexpr (PRE) {
                kbfv_tw pre[(1 << LR1) - 1];
                round_twiddles<LR1, LOG_S, COLUMN>(pre, LR0, tab, log_n, g0, grp0, stride);
                stage_in<LOG_S, COLUMN>(tile, in, grp0, stride);
                // pinned here so that the loads are not sunk below the barrier
#pragma unroll
                for (int i = 0; i < (1 << LR1) - 1; ++i) asm volatile("" : "+v"(pre[i].w), "+v"(pre[i].ws));
                __syncthreads();
                ntt_round<LR1, LOG_S, COLUMN, false, NI, false, false, false, true>(
                    tile, LR0, tab, log_n, g0, grp0, stride, q, ninv, nullptr, nullptr, &pre);
            } else {
                stage_in<LOG_S, COLUMN>(tile, in, grp0, stride);
                __syncthreads();
                ntt_round<LR1, LOG_S, COLUMN, false, NI>(tile, LR0, tab, log_n, g0, grp0, stride, q, ninv);
            }
        }
        ntt_round<LR0, LOG_S, COLUMN, false, NI, false, true, LAST && !NOSCALE, false, NOSCALE>(
            tile, 0, tab, log_n, g0, grp0, stride, q, ninv, nullptr, out);
        return;
    }

    // forward row pass: stride-1 round done in the tile, coalesced store from it
#pragma unroll
    for (int j = 0; j < TILE / T; ++j) {
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

template <int LOG_N, bool COLUMN, bool FWD, bool NOSCALE = false>
int launch_pass(const kbfv_context *c, const uint64_t *src, uint64_t *dst, int mod0, int limbs,
                int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL,
                hipStream_t st) {
    using G = Geo<pass_log<LOG_N, COLUMN>(), COLUMN>;
    const long groups = (1L << (LOG_N - pass_log<LOG_N, COLUMN>())) / G::C;
    const long blocks = groups * limbs * polys * batch;
    KB_REQUIRE(blocks < (1L << 31), "NTT grid too large");
    hipLaunchKernelGGL((k_ntt_pass<LOG_N, COLUMN, FWD, NOSCALE>), dim3((unsigned)blocks), dim3(G::T), 0, st, src, dst,
                       c->t.tw, c->t.ninv, c->t.inv_w1, c->t.mod, mod0, limbs, polys, batch, sB, sY, sL, dB, dY,
                       dL);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

template <bool COLUMN, bool FWD, bool NOSCALE = false>
int launch_pass_n(const kbfv_context *c, const uint64_t *src, uint64_t *dst, int mod0, int limbs, int polys,
                  int batch, long sB, long sY, long sL, long dB, long dY, long dL, hipStream_t st) {
#define KB_NTT_N(LN)                                                                                     \
    case LN:                                                                                             \
        return launch_pass<LN, COLUMN, FWD, NOSCALE>(c, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, \
                                            dL, st);
    switch (c->log_n) {
        KB_NTT_N(12)
        KB_NTT_N(13)
        KB_NTT_N(14)
        KB_NTT_N(15)
        KB_NTT_N(16)
        default:
            kb_set_error("unsupported ring degree 2^%d", c->log_n);
            return KBFV_EINVAL;
    }
#undef KB_NTT_N
}

}  // namespace

int kb_ntt(const kbfv_context *c, int dir, const uint64_t *src, uint64_t *dst, int mod0, int limbs,
           int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL,
           hipStream_t st) {
    KB_REQUIRE(c && src && dst, "null argument");
    KB_REQUIRE(c->log_n >= 12 && c->log_n <= 16, "NTT supports 2^12..2^16");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod, "modulus range out of bounds");
    KB_REQUIRE(polys > 0 && batch > 0, "empty NTT");
    if (mod0 + limbs - 1 == c->oPlain) KB_REQUIRE(c->batching, "plain modulus has no 2N-th root");
    int rc;
    if (dir > 0) {
        rc = launch_pass_n<true, true>(c, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, st);
        if (rc) return rc;
        return launch_pass_n<false, true>(c, dst, dst, mod0, limbs, polys, batch, dB, dY, dL, dB, dY, dL, st);
    }
    rc = launch_pass_n<false, false>(c, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, st);
    if (rc) return rc;
    return launch_pass_n<true, false>(c, dst, dst, mod0, limbs, polys, batch, dB, dY, dL, dB, dY,
                                      dL, st);
}

int kb_ntt_inplace(const kbfv_context *c, int dir, uint64_t *data, int mod0, int limbs, int polys,
                   hipStream_t st) {
    const long N = c->n;
    return kb_ntt(c, dir, data, data, mod0, limbs, polys, 1, 0, (long)limbs * N, N, 0,
                  (long)limbs * N, N, st);
}

int kb_intt_noscale(const kbfv_context *c, const uint64_t *src, uint64_t *dst, int mod0, int limbs, int polys,
                    int batch, long sB, long sY, long sL, long dB, long dY, long dL, hipStream_t st) {
    KB_REQUIRE(c && src && dst, "null argument");
    KB_REQUIRE(c->log_n >= 12 && c->log_n <= 16, "NTT supports 2^12..2^16");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod, "modulus range out of bounds");
    KB_REQUIRE(polys > 0 && batch > 0, "empty NTT");
    const int rc = launch_pass_n<false, false>(c, src, dst, mod0, limbs, polys, batch, sB, sY, sL, dB, dY, dL, st);
    if (rc) return rc;
    return launch_pass_n<true, false, true>(c, dst, dst, mod0, limbs, polys, batch, dB, dY, dL, dB, dY, dL, st);
}
