// Learned initial guess: a 3D U-Net b -> x0 for a fixed operator, trained on (b, x) pairs
// of the plain solver.  No reference counterpart (the reference always starts from the
// x the caller gives); the solver that runs from x0 is the unchanged one.
//
// Layout: activations channels-last (N, D, H, W, C), bf16; 3x3x3 convolutions with zero
// padding as implicit GEMMs on MFMA (v_mfma_f32_16x16x32_bf16, fp32 accumulation):
//   forward      out[m][co] = sum_{t,ci} in[m + off_t][ci] * W[t][ci][co]      (M x K . K x N)
//   data grad    gin[m][ci] = sum_{t,co} g[m + off_t][co] * W[26 - t][ci][co] (same kernel,
//                flipped and transposed weights)
//   weight grad  gW[t][ci][co] = sum_m in[m + off_t][ci] * g[m][co]          (LDS-staged
//                MFMA per (voxel chunk, tap), chunk partials reduced in fixed order)
// Weights are fp32 masters (Adam) with bf16 copies in the two GEMM layouts.
//
// MFMA 16x16x32 bf16 operand layout (CDNA3/CDNA4): lane l holds A[l % 16][8*(l/16) + j],
// B[8*(l/16) + j][l % 16] (j = 0..7) and D[4*(l/16) + i][l % 16] (i = 0..3);
// xh_unet_mfma_selftest checks it against a host product.
#include "../../../include/xamg_hip.h"
#include "../../../include/xamg_unet.h"
#include "xh_internal.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace xh {
namespace unet {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef unsigned short bf16_t; // storage

__device__ __forceinline__ float bf2f(bf16_t v) {
    return __uint_as_float((uint32_t)v << 16);
}

// round to nearest even
__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) // inf / nan: keep, quiet nan
        return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 pack8(const bf16_t (&e)[8]) {
    uint4 q;
    q.x = (uint32_t)e[0] | ((uint32_t)e[1] << 16);
    q.y = (uint32_t)e[2] | ((uint32_t)e[3] << 16);
    q.z = (uint32_t)e[4] | ((uint32_t)e[5] << 16);
    q.w = (uint32_t)e[6] | ((uint32_t)e[7] << 16);
    return __builtin_bit_cast(bf16x8, q);
}

struct Dims {
    int n, d, h, w; // batch and volume
    __host__ __device__ int64_t vox() const { return (int64_t)n * d * h * w; }
};

// neighbour of voxel m at tap t ((dz, dy, dx) = t / 9 - 1, (t / 3) % 3 - 1, t % 3 - 1),
// or -1 outside the volume
__device__ __forceinline__ int64_t neigh(const Dims &g, int64_t m, int t) {
    const int x = (int)(m % g.w), y = (int)((m / g.w) % g.h), z = (int)((m / ((int64_t)g.w * g.h)) % g.d);
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    if (zz < 0 || zz >= g.d || yy < 0 || yy >= g.h || xx < 0 || xx >= g.w)
        return -1;
    return m + ((int64_t)dz * g.h + dy) * g.w + dx;
}

// ---- convolution forward / data gradient: implicit GEMM on MFMA ----
// out[m][co] = act(sum_k A[m][k] B[k][co] + bias[co] + add[m][co]); A gathered from `in`
// (cin channels, k = t*cin + ci), B = wt[co][k] (kpad >= 27*cin, zero padded).
// 4 waves per workgroup; a wave computes MT tiles of 16 voxels for all cout/16 channel
// tiles, so every B fragment feeds MT MFMAs.  UP: `in` is the coarse volume of a nearest
// x2 upsampling (d/2 x h/2 x w/2) and the fine neighbour (z, y, x) reads coarse voxel
// (z/2, y/2, x/2): the upsampled tensor is never written.
template <int CT, int MT, bool UP>
__global__ __launch_bounds__(256) void k_conv(Dims g, const bf16_t *__restrict__ in, int cin,
                                              const bf16_t *__restrict__ wt, int kpad,
                                              const float *__restrict__ bias,
                                              const bf16_t *__restrict__ add, int relu,
                                              bf16_t *__restrict__ out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t M = g.vox();
    const int64_t m0 = ((int64_t)blockIdx.x * 4 + wave) * 16 * MT;
    if (m0 >= M)
        return;
    const int cout = CT * 16;
    const int K = 27 * cin;
    const int r = lane & 15, kg = lane >> 4;
    // the lane's voxel of each tile, decoded once
    int vx[MT], vy[MT], vz[MT], vb[MT];
    bool mok[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int64_t m = m0 + 16 * t + r;
        mok[t] = m < M;
        vx[t] = (int)(m % g.w);
        vy[t] = (int)((m / g.w) % g.h);
        const int64_t pl = m / ((int64_t)g.w * g.h);
        vz[t] = (int)(pl % g.d);
        vb[t] = (int)(pl / g.d);
    }
    f32x4 acc[MT][CT];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c)
            acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    // tap and channel of the lane's first k, advanced by 32 per step (cin % 8 == 0)
    int tt = (8 * kg) / cin, cc = 8 * kg - tt * cin;
    for (int kb = 0; kb < kpad; kb += 32) {
        const int k0 = kb + 8 * kg;
        bf16x8 a[MT];
        if (UP || (cin & 7) == 0) { // the fused upsampling is only used with cin % 8 == 0
            // 8 consecutive channels of one tap
            const int dz = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dx = tt % 3 - 1;
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                uint4 q = make_uint4(0, 0, 0, 0);
                const int zz = vz[t] + dz, yy = vy[t] + dy, xx = vx[t] + dx;
                if (mok[t] && k0 < K && zz >= 0 && zz < g.d && yy >= 0 && yy < g.h && xx >= 0 &&
                    xx < g.w) {
                    int64_t nb;
                    if constexpr (UP)
                        nb = (((int64_t)vb[t] * (g.d / 2) + zz / 2) * (g.h / 2) + yy / 2) *
                                 (g.w / 2) + xx / 2;
                    else
                        nb = m0 + 16 * t + r + ((int64_t)dz * g.h + dy) * g.w + dx;
                    q = *reinterpret_cast<const uint4 *>(in + nb * cin + cc);
                }
                a[t] = __builtin_bit_cast(bf16x8, q);
            }
            cc += 32;
            while (cc >= cin) {
                cc -= cin;
                ++tt;
            }
        } else {
            // cin not a multiple of 8 (the single-channel input layer): element gathers,
            // neighbours from the decoded voxel coordinates
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                bf16_t e[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = k0 + j;
                    bf16_t v = 0;
                    if (mok[t] && k < K) {
                        const int tp = k / cin, ci = k - tp * cin;
                        const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
                        const int zz = vz[t] + dz, yy = vy[t] + dy, xx = vx[t] + dx;
                        if (zz >= 0 && zz < g.d && yy >= 0 && yy < g.h && xx >= 0 && xx < g.w) {
                            const int64_t nb =
                                m0 + 16 * t + r + ((int64_t)dz * g.h + dy) * g.w + dx;
                            v = in[nb * cin + ci];
                        }
                    }
                    e[j] = v;
                }
                a[t] = pack8(e);
            }
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const uint4 qb = *reinterpret_cast<const uint4 *>(wt + (size_t)(c * 16 + r) * kpad + k0);
#pragma unroll
            for (int t = 0; t < MT; ++t)
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[t], __builtin_bit_cast(bf16x8, qb), acc[t][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int co = c * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t mo = m0 + 16 * t + 4 * kg + i;
                if (mo >= M)
                    continue;
                float v = acc[t][c][i] + (bias ? bias[co] : 0.f);
                if (add)
                    v = v + bf2f(add[mo * cout + co]);
                if (relu)
                    v = v > 0.f ? v : 0.f;
                out[mo * cout + co] = f2bf(v);
            }
        }
}

// ---- forward convolution over LDS-staged tiles (cout = 16 * CT) ----
// A workgroup computes kTZ x TY lines of TX voxels (64, or 32 for narrower grids).  The
// input those voxels read (the tile with a one-voxel halo, or with UP the coarse voxels
// under it) is staged in LDS once, zero outside the volume, each staged line stored as NCG
// groups of 8 channels (16 bytes) by voxel, so the 16 rows of an input fragment are 256
// contiguous bytes.  All 27 taps then read their input fragments from LDS instead of
// gathering them through the L1 per tap.  With CT = 1 and at most 16 k steps the weight
// fragments of all steps are loaded after the staging stores; otherwise each step's CT
// fragments are loaded one step ahead.  The weights are the A operand (the MFMA gives the
// transposed output tile, 4 channels of a voxel per lane, so the stores are 8-byte
// vectors); the products, their k order and the epilogue's arithmetic are those of
// k_conv, so the output is bit-identical.  The launch requires w % TX == 0, h % TY == 0
// and d % kTZ == 0.
constexpr int kTX = 64, kTY = 4, kTZ = 2;

// Epilogue of the tiled convolutions: lane (r, kg) of tile t, channel tile c holds
// channels 16c + 4kg .. 16c + 4kg + 3 of voxel r, written (and `add` read) as 8 bytes.
// Per element act(acc + bias + add), as in k_conv.
template <int CT, int NT, int XT, int LPW, int TY>
__device__ __forceinline__ void conv_tile_store(const Dims &g, int nb, int z0, int y0, int x0,
                                                int wave, int r, int kg,
                                                const f32x4 (&acc)[NT][CT],
                                                const float *__restrict__ bias,
                                                const bf16_t *__restrict__ add, int relu,
                                                bf16_t *__restrict__ out) {
    constexpr int COUT = 16 * CT;
    float bv[CT][4];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            bv[c][i] = bias ? bias[16 * c + 4 * kg + i] : 0.f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int li = wave * LPW + t / XT, xt = t % XT;
        const int iz = li / TY, iy = li % TY;
        const int64_t mo =
            (((int64_t)nb * g.d + z0 + iz) * g.h + y0 + iy) * g.w + x0 + 16 * xt + r;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int64_t o0 = mo * COUT + 16 * c + 4 * kg;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                v[i] = acc[t][c][i] + bv[c][i];
            if (add) {
                const uint2 a = *reinterpret_cast<const uint2 *>(add + o0);
                v[0] = v[0] + bf2f((bf16_t)(a.x & 0xffffu));
                v[1] = v[1] + bf2f((bf16_t)(a.x >> 16));
                v[2] = v[2] + bf2f((bf16_t)(a.y & 0xffffu));
                v[3] = v[3] + bf2f((bf16_t)(a.y >> 16));
            }
            if (relu)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    v[i] = v[i] > 0.f ? v[i] : 0.f;
            uint2 o;
            o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
            o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
            *reinterpret_cast<uint2 *>(out + o0) = o;
        }
    }
}

template <int NCG, int CT, int TY, bool UP, int TX = kTX>
__global__ __launch_bounds__(256) void k_conv_tile(Dims g, const bf16_t *__restrict__ in,
                                                   const bf16_t *__restrict__ wt, int kpad,
                                                   const float *__restrict__ bias,
                                                   const bf16_t *__restrict__ add, int relu,
                                                   bf16_t *__restrict__ out) {
    constexpr int CIN = 8 * NCG;
    constexpr int SZ = UP ? kTZ / 2 + 2 : kTZ + 2;
    constexpr int SY = UP ? TY / 2 + 2 : TY + 2;
    constexpr int SX = UP ? TX / 2 + 2 : TX + 2;
    constexpr int NCH = SZ * SY * SX * NCG; // 16-byte chunks staged
    constexpr int PER = (NCH + 255) / 256;
    constexpr int XT = TX / 16;               // 16-voxel tiles per line
    constexpr int LPW = kTZ * TY / 4;         // lines per wave
    constexpr int NT = LPW * XT;              // tiles per wave
    constexpr int KS = (27 * CIN + 31) / 32; // k steps (kpad == 32 * KS)
    constexpr bool PRE = CT == 1 && KS <= 16; // all weight fragments preloaded
    static_assert(LPW >= 1 && (kTZ * TY) % 4 == 0, "tile lines must split over 4 waves");
    __shared__ uint4 sm[NCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntx = g.w / TX, nty = g.h / TY, ntz = g.d / kTZ;
    int64_t bid = blockIdx.x;
    const int tx = (int)(bid % ntx);
    bid /= ntx;
    const int ty = (int)(bid % nty);
    bid /= nty;
    const int tz = (int)(bid % ntz);
    const int nb = (int)(bid / ntz);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * kTZ;
    const int r = lane & 15, kg = lane >> 4;
    // lane's weight row (output channel 16c + r of channel tile c at + 16c * kpad)
    const bf16_t *wrow = wt + (size_t)r * kpad + 8 * kg;
    // the staged volume (coarse with UP) and the tile's origin in it
    const int sd = UP ? g.d / 2 : g.d, sh = UP ? g.h / 2 : g.h, sw = UP ? g.w / 2 : g.w;
    const int oz = (UP ? z0 / 2 : z0) - 1, oy = (UP ? y0 / 2 : y0) - 1,
              ox = (UP ? x0 / 2 : x0) - 1;
    {
        uint4 st[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int q = tid + 256 * j;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (q < NCH) {
                const int line = q / (SX * NCG), rem = q - line * (SX * NCG);
                const int sx = rem / NCG, cg = rem - sx * NCG;
                const int zz = oz + line / SY, yy = oy + line % SY, xx = ox + sx;
                if (zz >= 0 && zz < sd && yy >= 0 && yy < sh && xx >= 0 && xx < sw)
                    v = *reinterpret_cast<const uint4 *>(
                        in + ((((int64_t)nb * sd + zz) * sh + yy) * sw + xx) * CIN + 8 * cg);
            }
            st[j] = v;
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int q = tid + 256 * j;
            if (q < NCH) {
                const int line = q / (SX * NCG), rem = q - line * (SX * NCG);
                const int sx = rem / NCG, cg = rem - sx * NCG;
                sm[(line * NCG + cg) * SX + sx] = st[j];
            }
        }
    }
    // weight fragments, issued once the staged values have left the registers
    uint4 qa[PRE ? KS : 1];
    if constexpr (PRE) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            qa[ks] = *reinterpret_cast<const uint4 *>(wrow + 32 * ks);
    }
    uint4 qn[CT];
    if constexpr (!PRE) {
#pragma unroll
        for (int c = 0; c < CT; ++c)
            qn[c] = *reinterpret_cast<const uint4 *>(wrow + (size_t)16 * c * kpad);
    }
    __syncthreads();
    f32x4 acc[NT][CT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c)
            acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int K = 27 * CIN;
    // tap and channel of the lane's first k, advanced by 32 per step (as in k_conv)
    int tt = (8 * kg) / CIN, cc = 8 * kg - tt * CIN;
    auto step = [&](int ks, const uint4 (&qb)[CT]) {
        const bool kok = 32 * ks + 8 * kg < K;
        const int dz = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dx = tt % 3 - 1;
        const int cg = cc >> 3;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int li = wave * LPW + t / XT, xt = t % XT;
            const int iz = li / TY, iy = li % TY;
            int sz, sy, sx;
            if constexpr (UP) {
                // fine offset -1 maps to coarse -1 (staged zero), as the fine bounds test
                sz = ((iz + dz) >> 1) + 1;
                sy = ((iy + dy) >> 1) + 1;
                sx = ((16 * xt + r + dx) >> 1) + 1;
            } else {
                sz = iz + dz + 1;
                sy = iy + dy + 1;
                sx = 16 * xt + r + dx + 1;
            }
            uint4 q = sm[kok ? ((sz * SY + sy) * NCG + cg) * SX + sx : 0];
            if (!kok)
                q = make_uint4(0, 0, 0, 0);
            // weights as the A operand: D[co][voxel], each lane 4 channels of one voxel
#pragma unroll
            for (int c = 0; c < CT; ++c)
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, qb[c]), __builtin_bit_cast(bf16x8, q), acc[t][c],
                    0, 0, 0);
        }
        cc += 32;
        while (cc >= CIN) {
            cc -= CIN;
            ++tt;
        }
    };
    if constexpr (PRE) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const uint4 qb[1] = {qa[ks]};
            step(ks, qb);
        }
    } else {
        for (int ks = 0; ks < KS; ++ks) {
            uint4 qb[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c)
                qb[c] = qn[c];
            if (ks + 1 < KS)
#pragma unroll
                for (int c = 0; c < CT; ++c)
                    qn[c] = *reinterpret_cast<const uint4 *>(wrow + (size_t)16 * c * kpad +
                                                             32 * (ks + 1));
            step(ks, qb);
        }
    }
    conv_tile_store<CT, NT, XT, LPW, TY>(g, nb, z0, y0, x0, wave, r, kg, acc, bias, add, relu,
                                         out);
}

// The single-channel input layer (cin = 1, K = 27 taps in one k step of 32) over the same
// tiles: the staged line holds one bf16 per voxel and each lane assembles its 8 taps of
// the A fragment from LDS, as k_conv assembles them from element gathers.
__global__ __launch_bounds__(256) void k_conv_tile1(Dims g, const bf16_t *__restrict__ in,
                                                    const bf16_t *__restrict__ wt, int kpad,
                                                    const float *__restrict__ bias, int relu,
                                                    bf16_t *__restrict__ out) {
    constexpr int SZ = kTZ + 2, SY = kTY + 2, SX = kTX + 2;
    constexpr int NS = SZ * SY * SX;
    constexpr int PER = (NS + 255) / 256;
    constexpr int XT = kTX / 16, LPW = kTZ * kTY / 4, NT = LPW * XT;
    __shared__ bf16_t sm[NS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntx = g.w / kTX, nty = g.h / kTY, ntz = g.d / kTZ;
    int64_t bid = blockIdx.x;
    const int tx = (int)(bid % ntx);
    bid /= ntx;
    const int ty = (int)(bid % nty);
    bid /= nty;
    const int tz = (int)(bid % ntz);
    const int nb = (int)(bid / ntz);
    const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
    {
        bf16_t st[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int q = tid + 256 * j;
            bf16_t v = 0;
            if (q < NS) {
                const int line = q / SX, sx = q - line * SX;
                const int zz = z0 - 1 + line / SY, yy = y0 - 1 + line % SY, xx = x0 - 1 + sx;
                if (zz >= 0 && zz < g.d && yy >= 0 && yy < g.h && xx >= 0 && xx < g.w)
                    v = in[(((int64_t)nb * g.d + zz) * g.h + yy) * g.w + xx];
            }
            st[j] = v;
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int q = tid + 256 * j;
            if (q < NS)
                sm[q] = st[j];
        }
    }
    __syncthreads();
    const int r = lane & 15, kg = lane >> 4;
    const uint4 qb = *reinterpret_cast<const uint4 *>(wt + (size_t)r * kpad + 8 * kg);
    f32x4 acc[NT][1];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int li = wave * LPW + t / XT, xt = t % XT;
        const int iz = li / kTY, iy = li % kTY;
        bf16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int tp = 8 * kg + j;
            const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
            const int idx = ((iz + dz + 1) * SY + iy + dy + 1) * SX + 16 * xt + r + dx + 1;
            e[j] = tp < 27 ? sm[tp < 27 ? idx : 0] : (bf16_t)0;
        }
        acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, qb), pack8(e), f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    conv_tile_store<1, NT, XT, LPW, kTY>(g, nb, z0, y0, x0, wave, r, kg, acc, bias, nullptr,
                                         relu, out);
}

// ---- weight gradient: per (voxel chunk, tap) workgroup, LDS-staged MFMA ----
// part[chunk][t][ci][co] = sum_{m in chunk} in[m + off_t][ci] * g[m][co]
// (ci padded to a multiple of 16 with zero rows; cin <= 64, cout <= 64)
__global__ __launch_bounds__(256) void k_wgrad(Dims gd, const bf16_t *__restrict__ in, int cin,
                                               const bf16_t *__restrict__ gr, int cout,
                                               int64_t chunk, float *__restrict__ part) {
    __shared__ bf16_t sa[32][64 + 8]; // [voxel][ci]
    __shared__ bf16_t sb[32][64 + 8]; // [voxel][co]
    const int t = blockIdx.y;
    const int64_t c0 = (int64_t)blockIdx.x * chunk;
    const int64_t M = gd.vox();
    const int64_t c1 = std::min<int64_t>(M, c0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kg = lane >> 4;
    const int rt_n = (cin + 15) / 16, ct_n = cout / 16, ntile = rt_n * ct_n;
    f32x4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t v0 = c0; v0 < c1; v0 += 32) {
        // stage 32 voxels: in at the tap neighbour, g at the voxel
        for (int e = threadIdx.x; e < 32 * 64; e += 256) {
            const int vv = e / 64, c = e % 64;
            const int64_t m = v0 + vv;
            bf16_t va = 0, vb = 0;
            if (m < c1) {
                if (c < cin) {
                    const int64_t nb = neigh(gd, m, t);
                    if (nb >= 0)
                        va = in[nb * cin + c];
                }
                if (c < cout)
                    vb = gr[m * cout + c];
            }
            sa[vv][c] = va;
            sb[vv][c] = vb;
        }
        __syncthreads();
        for (int q = 0; q < 4; ++q) {
            const int tile = wave + 4 * q;
            if (tile >= ntile)
                break;
            const int rt = tile / ct_n, ct = tile % ct_n;
            bf16_t ea[8], eb[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                ea[j] = sa[8 * kg + j][rt * 16 + r]; // A[ci][voxel]
                eb[j] = sb[8 * kg + j][ct * 16 + r]; // B[voxel][co]
            }
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack8(ea), pack8(eb), acc[q], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int q = 0; q < 4; ++q) {
        const int tile = wave + 4 * q;
        if (tile >= ntile)
            break;
        const int rt = tile / ct_n, ct = tile % ct_n;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int ci = rt * 16 + 4 * kg + i, co = ct * 16 + r;
            if (ci < cin)
                part[(((size_t)blockIdx.x * 27 + t) * cin + ci) * cout + co] = acc[q][i];
        }
    }
}

// out[i] = sum_c part[c][i], c = 0..nchunk-1 in order
__global__ void k_reduce_parts(const float *__restrict__ part, int nchunk, int64_t len,
                               float *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nchunk; ++c)
            s = s + part[(size_t)c * len + i];
        out[i] = s;
    }
}

// per-chunk channel sums of a [M][C] tensor (bias gradient): part[chunk][c]
__global__ void k_chansum(const bf16_t *__restrict__ g, int64_t M, int C, int64_t chunk,
                          float *__restrict__ part) {
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = std::min<int64_t>(M, c0 + chunk);
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        float s = 0.f;
        for (int64_t m = c0; m < c1; ++m)
            s = s + bf2f(g[m * C + c]);
        part[(size_t)blockIdx.x * C + c] = s;
    }
}

// ---- elementwise layers ----
__global__ void k_relu_mask(int64_t n, const bf16_t *__restrict__ act, bf16_t *__restrict__ g) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (!(bf2f(act[i]) > 0.f))
            g[i] = 0;
}

// a += b
__global__ void k_add(int64_t n, bf16_t *__restrict__ a, const bf16_t *__restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// 2x2x2 average pooling (fine dims g, coarse g/2), C channels
__global__ void k_pool(Dims gf, int C, const bf16_t *__restrict__ in, bf16_t *__restrict__ out) {
    const int D = gf.d / 2, H = gf.h / 2, W = gf.w / 2;
    const int64_t n = (int64_t)gf.n * D * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t v = i / C;
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)((v / ((int64_t)W * H)) % D);
        const int b = (int)(v / ((int64_t)W * H * D));
        float s = 0.f;
        for (int q = 0; q < 8; ++q) {
            const int64_t f = (((int64_t)b * gf.d + 2 * z + (q >> 2)) * gf.h + 2 * y + ((q >> 1) & 1)) *
                                  gf.w + 2 * x + (q & 1);
            s = s + bf2f(in[f * C + c]);
        }
        out[i] = f2bf(s * 0.125f);
    }
}

// gradient of the pooling: every fine cell gets g_coarse / 8
__global__ void k_pool_bwd(Dims gf, int C, const bf16_t *__restrict__ gc, bf16_t *__restrict__ gfine) {
    const int64_t n = gf.vox() * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t v = i / C;
        const int x = (int)(v % gf.w), y = (int)((v / gf.w) % gf.h), z = (int)((v / ((int64_t)gf.w * gf.h)) % gf.d);
        const int b = (int)(v / ((int64_t)gf.w * gf.h * gf.d));
        const int64_t cv = (((int64_t)b * (gf.d / 2) + z / 2) * (gf.h / 2) + y / 2) * (gf.w / 2) + x / 2;
        gfine[i] = f2bf(bf2f(gc[cv * C + c]) * 0.125f);
    }
}

// nearest upsampling x2 (coarse -> fine dims gf)
__global__ void k_up(Dims gf, int C, const bf16_t *__restrict__ in, bf16_t *__restrict__ out) {
    const int64_t n = gf.vox() * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t v = i / C;
        const int x = (int)(v % gf.w), y = (int)((v / gf.w) % gf.h), z = (int)((v / ((int64_t)gf.w * gf.h)) % gf.d);
        const int b = (int)(v / ((int64_t)gf.w * gf.h * gf.d));
        const int64_t cv = (((int64_t)b * (gf.d / 2) + z / 2) * (gf.h / 2) + y / 2) * (gf.w / 2) + x / 2;
        out[i] = in[cv * C + c];
    }
}

// gradient of the upsampling: coarse cell gets the sum of its 8 fine cells
__global__ void k_up_bwd(Dims gf, int C, const bf16_t *__restrict__ gfine, bf16_t *__restrict__ gc) {
    const int D = gf.d / 2, H = gf.h / 2, W = gf.w / 2;
    const int64_t n = (int64_t)gf.n * D * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t v = i / C;
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)((v / ((int64_t)W * H)) % D);
        const int b = (int)(v / ((int64_t)W * H * D));
        float s = 0.f;
        for (int q = 0; q < 8; ++q) {
            const int64_t f = (((int64_t)b * gf.d + 2 * z + (q >> 2)) * gf.h + 2 * y + ((q >> 1) & 1)) *
                                  gf.w + 2 * x + (q & 1);
            s = s + bf2f(gfine[f * C + c]);
        }
        gc[i] = f2bf(s);
    }
}

// 1x1 head: y[m] = sum_c w[c] h[m][c] + b
__global__ void k_head(int64_t M, int C, const bf16_t *__restrict__ h, const float *__restrict__ w,
                       const float *__restrict__ b, float *__restrict__ y) {
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M;
         m += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < C; ++c)
            s = s + w[c] * bf2f(h[m * C + c]);
        y[m] = s + b[0];
    }
}

// inference head fused with the output scaling: x0[v*nv + k] = s_k * o * (w . h[k*V + v] + b),
// threads ordered by (v, k) with k fastest so x0 is written contiguously (each thread
// reads its voxel's C channels, one 16-byte-aligned run)
__global__ void k_head_out(int64_t V, int nv, int C, const bf16_t *__restrict__ h,
                           const float *__restrict__ w, const float *__restrict__ b,
                           const float *__restrict__ scale, float o, double *__restrict__ x0) {
    const int64_t n = V * nv;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % nv);
        const int64_t v = i / nv, m = (int64_t)k * V + v;
        float s = 0.f;
        for (int c = 0; c < C; ++c)
            s = s + w[c] * bf2f(h[m * C + c]);
        const float y = s + b[0];
        x0[i] = (double)scale[k] * (double)o * (double)y;
    }
}

// the same outputs for nv | 256 and C % 8 == 0: a block writes 256 consecutive outputs
// (256 / nv voxels, all nv RHS); each thread computes one (voxel, RHS) from its C channels
// read as 16-byte vectors (the block's reads of one sample are contiguous) and the block
// writes its outputs in order through LDS.  Per-output arithmetic as k_head_out.
__global__ __launch_bounds__(256) void k_head_out_t(int64_t V, int nv, int C,
                                                    const bf16_t *__restrict__ h,
                                                    const float *__restrict__ w,
                                                    const float *__restrict__ b,
                                                    const float *__restrict__ scale, float o,
                                                    double *__restrict__ x0) {
    __shared__ double sm[256];
    const int t = threadIdx.x, vb = 256 / nv;
    const int64_t i0 = (int64_t)blockIdx.x * 256, v0 = i0 / nv;
    const int k = t / vb, dv = t - k * vb;
    const int64_t v = v0 + dv;
    if (v < V) {
        const uint4 *hp = reinterpret_cast<const uint4 *>(h + ((int64_t)k * V + v) * C);
        float s = 0.f;
        for (int c8 = 0; c8 < C / 8; ++c8) {
            const uint4 q = hp[c8];
            const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s = s + w[8 * c8 + 2 * j] * bf2f((bf16_t)(u[j] & 0xffffu));
                s = s + w[8 * c8 + 2 * j + 1] * bf2f((bf16_t)(u[j] >> 16));
            }
        }
        const float y = s + b[0];
        sm[dv * nv + k] = (double)scale[k] * (double)o * (double)y;
    }
    __syncthreads();
    if (i0 + t < V * nv)
        x0[i0 + t] = sm[t];
}

// head gradients: gh[m][c] = gy[m] w[c]; per-chunk partials of gw[c] = sum gy h, gb = sum gy
__global__ void k_head_bwd(int64_t M, int C, const bf16_t *__restrict__ h, const float *__restrict__ w,
                           const float *__restrict__ gy, bf16_t *__restrict__ gh, int64_t chunk,
                           float *__restrict__ part) {
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = std::min<int64_t>(M, c0 + chunk);
    for (int64_t m = c0 + threadIdx.x; m < c1; m += blockDim.x)
        for (int c = 0; c < C; ++c)
            gh[m * C + c] = f2bf(gy[m] * w[c]);
    for (int c = threadIdx.x; c <= C; c += blockDim.x) {
        float s = 0.f;
        for (int64_t m = c0; m < c1; ++m)
            s = s + gy[m] * (c < C ? bf2f(h[m * C + c]) : 1.f);
        part[(size_t)blockIdx.x * (C + 1) + c] = s;
    }
}

// loss partials: 0.5 (y - t)^2 * inv_m summed per chunk; gy = (y - t) * inv_m
__global__ void k_loss(int64_t M, const float *__restrict__ y, const float *__restrict__ tgt,
                       float inv_m, float *__restrict__ gy, int64_t chunk, float *__restrict__ part) {
    __shared__ float red[256];
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = std::min<int64_t>(M, c0 + chunk);
    float s = 0.f;
    for (int64_t m = c0 + threadIdx.x; m < c1; m += blockDim.x) {
        const float d = y[m] - tgt[m];
        gy[m] = d * inv_m;
        s = s + 0.5f * d * d * inv_m;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int i = 0; i < (int)blockDim.x; ++i)
            t = t + red[i];
        part[blockIdx.x] = t;
    }
}

// bf16 GEMM layouts of a 3x3x3 layer from its fp32 masters w[t][ci][co]:
//   wt [co][kpad]  (k = t*cin + ci)       forward
//   wtf[ci][kpadf] (k = t*cout + co) = w[26 - t][ci][co]   data gradient
__global__ void k_prep(const float *__restrict__ w, int cin, int cout, int kpad, int kpadf,
                       bf16_t *__restrict__ wt, bf16_t *__restrict__ wtf) {
    const int64_t n1 = (int64_t)cout * kpad, n2 = (int64_t)cin * kpadf;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n1) {
            const int co = (int)(i / kpad), k = (int)(i % kpad);
            float v = 0.f;
            if (k < 27 * cin) {
                const int t = k / cin, ci = k % cin;
                v = w[((size_t)t * cin + ci) * cout + co];
            }
            wt[i] = f2bf(v);
        } else {
            const int64_t j = i - n1;
            const int ci = (int)(j / kpadf), k = (int)(j % kpadf);
            float v = 0.f;
            if (k < 27 * cout) {
                const int t = k / cout, co = k % cout;
                v = w[((size_t)(26 - t) * cin + ci) * cout + co];
            }
            wtf[j] = f2bf(v);
        }
    }
}

// Adam (Kingma & Ba): m = b1 m + (1-b1) g ; v = b2 v + (1-b2) g^2 ;
// p -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps)
__global__ void k_adam(int64_t n, float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// per-sample RMS of an fp64 [vox][nv] multivector slice: scale[k] (partials per chunk)
// per block: rows [c0, c1) of the interleaved [V][nv] multivector; thread t sums column
// k = t % nv over the rows c0 + q, c0 + q + RL, ... (q = t / nv, RL = 256 / nv phases) and
// the block partial adds the phase sums in the order q = 0, 1, ... (nv <= 64)
__global__ __launch_bounds__(256) void k_rms_part(int64_t V, int nv, const double *__restrict__ b,
                                                  int64_t chunk, double *__restrict__ part) {
    __shared__ double red[256];
    const int t = threadIdx.x, RL = 256 / nv, k = t % nv, q = t / nv;
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = std::min<int64_t>(V, c0 + chunk);
    double s = 0.0;
    if (q < RL) {
#pragma unroll 8
        for (int64_t i = c0 + q; i < c1; i += RL) {
            const double x = b[i * nv + k];
            s = s + x * x;
        }
    }
    red[t] = s;
    __syncthreads();
    if (t < nv) {
        double a = 0.0;
        for (int j = 0; j < RL; ++j)
            a = a + red[j * nv + t];
        part[(size_t)blockIdx.x * nv + t] = a;
    }
}

// scale[k] = rms of column k: the chunk partials summed in the same phase order
__global__ __launch_bounds__(256) void k_rms_fin(int nchunk, int nv, int64_t V,
                                                 const double *__restrict__ part,
                                                 float *__restrict__ scale) {
    __shared__ double red[256];
    const int t = threadIdx.x, RL = 256 / nv, k = t % nv, q = t / nv;
    double s = 0.0;
    if (q < RL)
        for (int c = q; c < nchunk; c += RL)
            s = s + part[(size_t)c * nv + k];
    red[t] = s;
    __syncthreads();
    if (t < nv) {
        double a = 0.0;
        for (int j = 0; j < RL; ++j)
            a = a + red[j * nv + t];
        const double r = sqrt(a / (double)V);
        scale[t] = r > 0.0 ? (float)r : 1.f;
    }
}

// network input (sample k = batch index) from column k of an interleaved [vox][nv] fp64
// multivector: in[k][vox] = b[vox][k] / scale[k] (bf16, 1 channel); tgt likewise (fp32).
// A block transposes 64 voxels x nv (nv <= 64) through LDS: the multivector rows are read
// contiguously and each sample's 64 voxels are written contiguously.
__global__ __launch_bounds__(256) void k_gather_in(int64_t V, int nv, const double *__restrict__ b,
                                                   const float *__restrict__ scale,
                                                   bf16_t *__restrict__ in,
                                                   const double *__restrict__ x,
                                                   float *__restrict__ tgt) {
    constexpr int P = 65; // padded LDS row
    __shared__ bf16_t sb[64 * P];
    __shared__ float sx[64 * P];
    const int t = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int nvox = (int)std::min<int64_t>(64, V - v0);
    for (int e = t; e < nvox * nv; e += 256) { // e = dv * nv + k, contiguous in b
        const int dv = e / nv, k = e - dv * nv;
        const size_t gi = (size_t)v0 * nv + e;
        sb[k * P + dv] = f2bf((float)(b[gi] / (double)scale[k]));
        if (tgt)
            sx[k * P + dv] = (float)(x[gi] / (double)scale[k]);
    }
    __syncthreads();
    for (int e = t; e < nv * 64; e += 256) { // e = k * 64 + dv, contiguous in `in`
        const int k = e >> 6, dv = e & 63;
        if (dv < nvox) {
            in[(size_t)k * V + v0 + dv] = sb[k * P + dv];
            if (tgt)
                tgt[(size_t)k * V + v0 + dv] = sx[k * P + dv];
        }
    }
}


// a[i] *= f
__global__ void k_scale_f32(int64_t n, float *__restrict__ a, float f) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = a[i] * f;
}

// crop of a [nv] batch of full volumes into a batch of sub-volumes (bf16 input, fp32 target)
__global__ void k_crop(Dims gfull, Dims gcrop, const int *__restrict__ org,
                       const bf16_t *__restrict__ in, const float *__restrict__ tgt,
                       bf16_t *__restrict__ cin, float *__restrict__ ctgt) {
    const int64_t n = gcrop.vox();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % gcrop.w), y = (int)((i / gcrop.w) % gcrop.h),
                  z = (int)((i / ((int64_t)gcrop.w * gcrop.h)) % gcrop.d);
        const int b = (int)(i / ((int64_t)gcrop.w * gcrop.h * gcrop.d));
        const int s = org[4 * b];
        const int64_t f = (((int64_t)s * gfull.d + org[4 * b + 1] + z) * gfull.h + org[4 * b + 2] + y) *
                              gfull.w + org[4 * b + 3] + x;
        cin[i] = in[f];
        ctgt[i] = tgt[f];
    }
}

__global__ void k_f2bf(int64_t n, const float *__restrict__ a, bf16_t *__restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

__global__ void k_mfma_selftest(const float *__restrict__ a, const float *__restrict__ b,
                                float *__restrict__ d) {
    // a: 16x32 row-major, b: 32x16 row-major, d: 16x16 row-major
    const int l = threadIdx.x, r = l & 15, kg = l >> 4;
    bf16_t ea[8], eb[8];
    for (int j = 0; j < 8; ++j) {
        ea[j] = f2bf(a[r * 32 + 8 * kg + j]);
        eb[j] = f2bf(b[(8 * kg + j) * 16 + r]);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack8(ea), pack8(eb), acc, 0, 0, 0);
    for (int i = 0; i < 4; ++i)
        d[(4 * kg + i) * 16 + r] = acc[i];
}

} // namespace unet
} // namespace xh

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
namespace xh {
namespace unet {
namespace {

#define UN_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return report_error(XH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

int grid1(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

int64_t rup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unif(uint64_t seed, uint64_t i) {
    return (double)(mix64(seed * 0xD1B54A32D192ED03ull + i) >> 11) * 0x1.0p-53;
}

double normal(uint64_t seed, uint64_t i) {
    const double u1 = 1.0 - unif(seed, 2 * i), u2 = unif(seed, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Conv {
    int cin = 0, cout = 0, kpad = 0, kpadf = 0;
    int64_t nw = 0;
    float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    bf16_t *wt = nullptr, *wtf = nullptr;
};

// activations (and gradients when training) of one batch shape; level l has dims
// (n, d >> l, h >> l, w >> l)
struct Work {
    int n = 0, d = 0, h = 0, w = 0;
    bool grads = false;
    std::vector<Dims> g;
    std::vector<bf16_t *> in, a, s, u, dd, hh;       // u, dd, hh for l < L-1
    std::vector<bf16_t *> gin, ga, gs, gu, gdd, ghh, tmp;
    float *y = nullptr, *tgt = nullptr, *gy = nullptr;
    std::vector<void *> owned;
};

} // namespace
} // namespace unet
} // namespace xh

struct xh_unet_s {
    int L = 0;
    std::vector<int> ch;
    std::vector<xh::unet::Conv> enc, dec; // enc[2l] A, enc[2l+1] B; dec[2l] U, dec[2l+1] D
    int hc = 0;                           // head channels (c0)
    float *hw = nullptr, *hb = nullptr, *hgw = nullptr, *hgb = nullptr;
    float *hmw = nullptr, *hvw = nullptr, *hmb = nullptr, *hvb = nullptr;
    float oscale = 1.f;
    int64_t nparams = 0;
    int adam_t = 0;
    float *part = nullptr;  // chunk partials
    size_t part_len = 0;
    float *scale = nullptr; // per-sample RMS
    float *dloss = nullptr; // loss of the last forward pass
    double *dpart = nullptr;
    std::vector<std::unique_ptr<xh::unet::Work>> work;
    std::vector<void *> owned;
    hipStream_t stream = nullptr;
};

namespace xh {
namespace unet {
namespace {

int dalloc(xh_unet_s *U, void **p, size_t bytes, std::vector<void *> &own) {
    (void)U;
    UN_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
    UN_HIP(hipMemset(*p, 0, std::max<size_t>(bytes, 16)));
    own.push_back(*p);
    return 0;
}

template <typename T>
int dalloc_t(xh_unet_s *U, T **p, size_t count, std::vector<void *> &own) {
    return dalloc(U, (void **)p, count * sizeof(T), own);
}

int make_conv(xh_unet_s *U, int cin, int cout, uint64_t seed, uint64_t &draw, Conv &c) {
    c.cin = cin;
    c.cout = cout;
    c.kpad = (int)rup(27 * cin, 32);
    c.kpadf = (int)rup(27 * cout, 32);
    c.nw = 27LL * cin * cout;
    int rc;
    for (float **q : {&c.w, &c.gw, &c.mw, &c.vw})
        if ((rc = dalloc_t(U, q, c.nw, U->owned)))
            return rc;
    for (float **q : {&c.b, &c.gb, &c.mb, &c.vb})
        if ((rc = dalloc_t(U, q, cout, U->owned)))
            return rc;
    if ((rc = dalloc_t(U, &c.wt, (size_t)cout * c.kpad, U->owned)) ||
        (rc = dalloc_t(U, &c.wtf, (size_t)cin * c.kpadf, U->owned)))
        return rc;
    // He initialisation
    std::vector<float> hw(c.nw);
    const double sd = std::sqrt(2.0 / (27.0 * cin));
    for (int64_t i = 0; i < c.nw; ++i)
        hw[i] = (float)(sd * normal(seed, draw++));
    UN_HIP(hipMemcpy(c.w, hw.data(), c.nw * 4, hipMemcpyHostToDevice));
    return 0;
}

void prep(xh_unet_s *U, const Conv &c) {
    const int64_t n = (int64_t)c.cout * c.kpad + (int64_t)c.cin * c.kpadf;
    hipLaunchKernelGGL(k_prep, dim3(grid1(n)), dim3(256), 0, U->stream, c.w, c.cin, c.cout,
                       c.kpad, c.kpadf, c.wt, c.wtf);
}

void prep_all(xh_unet_s *U) {
    for (const Conv &c : U->enc)
        prep(U, c);
    for (const Conv &c : U->dec)
        prep(U, c);
}

int level_dims_ok(const xh_unet_s *U, int d, int h, int w) {
    const int f = 1 << (U->L - 1);
    return d > 0 && h > 0 && w > 0 && d % f == 0 && h % f == 0 && w % f == 0;
}

int get_work(xh_unet_s *U, int n, int d, int h, int w, bool grads, Work **out) {
    for (auto &p : U->work)
        if (p->n == n && p->d == d && p->h == h && p->w == w && (p->grads || !grads)) {
            *out = p.get();
            return 0;
        }
    auto W = std::make_unique<Work>();
    W->n = n;
    W->d = d;
    W->h = h;
    W->w = w;
    W->grads = grads;
    const int L = U->L;
    int rc;
    for (int l = 0; l < L; ++l) {
        const Dims g{n, d >> l, h >> l, w >> l};
        W->g.push_back(g);
        const int64_t V = g.vox();
        const int cl = U->ch[l], cinl = l == 0 ? 1 : U->ch[l - 1];
        const int cup = l + 1 < L ? U->ch[l + 1] : 0;
        bf16_t *p_in, *p_a, *p_s, *p_u = nullptr, *p_d = nullptr, *p_h = nullptr;
        if ((rc = dalloc_t(U, &p_in, V * cinl, W->owned)) || (rc = dalloc_t(U, &p_a, V * cl, W->owned)) ||
            (rc = dalloc_t(U, &p_s, V * cl, W->owned)))
            return rc;
        if (l + 1 < L)
            if ((rc = dalloc_t(U, &p_u, V * cup, W->owned)) || (rc = dalloc_t(U, &p_d, V * cl, W->owned)) ||
                (rc = dalloc_t(U, &p_h, V * cl, W->owned)))
                return rc;
        W->in.push_back(p_in);
        W->a.push_back(p_a);
        W->s.push_back(p_s);
        W->u.push_back(p_u);
        W->dd.push_back(p_d);
        W->hh.push_back(p_h);
        bf16_t *g_in = nullptr, *g_a = nullptr, *g_s = nullptr, *g_u = nullptr, *g_d = nullptr,
               *g_h = nullptr, *g_t = nullptr;
        if (grads) {
            if ((rc = dalloc_t(U, &g_in, V * cinl, W->owned)) || (rc = dalloc_t(U, &g_a, V * cl, W->owned)) ||
                (rc = dalloc_t(U, &g_s, V * cl, W->owned)) || (rc = dalloc_t(U, &g_t, V * cl, W->owned)))
                return rc;
            if (l + 1 < L)
                if ((rc = dalloc_t(U, &g_u, V * cup, W->owned)) ||
                    (rc = dalloc_t(U, &g_d, V * cl, W->owned)) || (rc = dalloc_t(U, &g_h, V * cl, W->owned)))
                    return rc;
        }
        W->gin.push_back(g_in);
        W->ga.push_back(g_a);
        W->gs.push_back(g_s);
        W->gu.push_back(g_u);
        W->gdd.push_back(g_d);
        W->ghh.push_back(g_h);
        W->tmp.push_back(g_t);
    }
    const int64_t V0 = W->g[0].vox();
    if ((rc = dalloc_t(U, &W->y, V0, W->owned)) || (rc = dalloc_t(U, &W->tgt, V0, W->owned)) ||
        (rc = dalloc_t(U, &W->gy, V0, W->owned)))
        return rc;
    *out = W.get();
    U->work.push_back(std::move(W));
    return 0;
}

// MT tiles of 16 voxels per wave: CT * MT = 4 accumulator tiles
template <int CT, bool UP = false>
void launch_conv(const Dims &g, const bf16_t *in, int cin, const bf16_t *wt, int kpad,
                 const float *bias, const bf16_t *add, int relu, bf16_t *out, hipStream_t s) {
    constexpr int MT = CT >= 3 ? 1 : 4 / CT;
    const dim3 grid((unsigned)((g.vox() + 64 * MT - 1) / (64 * MT))), blk(256);
    hipLaunchKernelGGL((k_conv<CT, MT, UP>), grid, blk, 0, s, g, in, cin, wt, kpad, bias, add,
                       relu, out);
}

// 0: every forward convolution runs as k_conv (no LDS-staged tiles)
int g_conv_tile = 1;

template <int NCG, int CT, int TY, bool UP, int TX = kTX>
void launch_conv_tile(const Dims &g, const bf16_t *in, const bf16_t *wt, int kpad,
                      const float *bias, const bf16_t *add, int relu, bf16_t *out,
                      hipStream_t s) {
    const dim3 grid((unsigned)(g.vox() / (TX * TY * kTZ))), blk(256);
    hipLaunchKernelGGL((k_conv_tile<NCG, CT, TY, UP, TX>), grid, blk, 0, s, g, in, wt, kpad,
                       bias, add, relu, out);
}

// up: `in` is the coarse input of a nearest x2 upsampling onto g (fused)
void conv_fwd(xh_unet_s *U, const Conv &c, const Dims &g, const bf16_t *in, const bf16_t *add,
              bool relu, bf16_t *out, bool up = false) {
    // 16 or 32 output channels on volumes the tiles divide: LDS-staged tiles (the staged
    // input at most 68 KB: 32 plain input channels take tiles of 2 lines)
    const bool tiles = g_conv_tile && g.w % kTX == 0 && g.h % kTY == 0 && g.d % kTZ == 0;
    hipStream_t st = U->stream;
    const int rl = (int)relu;
    if (tiles && c.cout == 16 && c.cin == 1 && !up && !add) {
        const dim3 grid((unsigned)(g.vox() / (kTX * kTY * kTZ))), blk(256);
        hipLaunchKernelGGL(k_conv_tile1, grid, blk, 0, st, g, in, c.wt, c.kpad, c.b, rl, out);
        return;
    }
    if (tiles && c.cout == 16) {
        if (c.cin == 16 && !up) {
            launch_conv_tile<2, 1, 4, false>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 16 && up) {
            launch_conv_tile<2, 1, 4, true>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 32 && up) {
            launch_conv_tile<4, 1, 4, true>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
    }
    if (tiles && c.cout == 32) {
        if (c.cin == 16 && !up) {
            launch_conv_tile<2, 2, 4, false>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 32 && !up) {
            launch_conv_tile<4, 2, 2, false>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 32 && up) {
            launch_conv_tile<4, 2, 4, true>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 48 && up) {
            launch_conv_tile<6, 2, 4, true>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
    }
    // 48 output channels on lines of 32 voxels (the third level of a 128^3 grid)
    if (g_conv_tile && c.cout == 48 && g.w % 32 == 0 && g.h % kTY == 0 && g.d % kTZ == 0) {
        if (c.cin == 32 && !up) {
            launch_conv_tile<4, 3, 4, false, 32>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 48 && !up) {
            launch_conv_tile<6, 3, 2, false, 32>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
        if (c.cin == 64 && up) {
            launch_conv_tile<8, 3, 4, true, 32>(g, in, c.wt, c.kpad, c.b, add, rl, out, st);
            return;
        }
    }
#define XH_CONV_CASE(CTV)                                                                      \
    if (up)                                                                                    \
        launch_conv<CTV, true>(g, in, c.cin, c.wt, c.kpad, c.b, add, (int)relu, out,          \
                               U->stream);                                                     \
    else                                                                                       \
        launch_conv<CTV, false>(g, in, c.cin, c.wt, c.kpad, c.b, add, (int)relu, out,         \
                                U->stream);
    switch (c.cout / 16) {
    case 1:
        XH_CONV_CASE(1)
        break;
    case 2:
        XH_CONV_CASE(2)
        break;
    case 3:
        XH_CONV_CASE(3)
        break;
    default:
        XH_CONV_CASE(4)
        break;
    }
#undef XH_CONV_CASE
}

// gin = data gradient of conv c (c.cin must be a multiple of 16)
void conv_data(xh_unet_s *U, const Conv &c, const Dims &g, const bf16_t *gout, bf16_t *gin) {
    switch (c.cin / 16) {
    case 1:
        launch_conv<1>(g, gout, c.cout, c.wtf, c.kpadf, nullptr, nullptr, 0, gin, U->stream);
        break;
    case 2:
        launch_conv<2>(g, gout, c.cout, c.wtf, c.kpadf, nullptr, nullptr, 0, gin, U->stream);
        break;
    case 3:
        launch_conv<3>(g, gout, c.cout, c.wtf, c.kpadf, nullptr, nullptr, 0, gin, U->stream);
        break;
    default:
        launch_conv<4>(g, gout, c.cout, c.wtf, c.kpadf, nullptr, nullptr, 0, gin, U->stream);
        break;
    }
}

constexpr int kMaxChunks = 256;

void conv_wgrad(xh_unet_s *U, const Conv &c, const Dims &g, const bf16_t *in, const bf16_t *gout) {
    const int64_t V = g.vox();
    const int nch = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxChunks, (V + 511) / 512));
    const int64_t chunk = rup((V + nch - 1) / nch, 32);
    const int nchunk = (int)((V + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_wgrad, dim3(nchunk, 27), dim3(256), 0, U->stream, g, in, c.cin, gout,
                       c.cout, chunk, U->part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid1(c.nw)), dim3(256), 0, U->stream, U->part,
                       nchunk, c.nw, c.gw);
    const int64_t cb = (V + 1023) / 1024;
    const int nb = (int)((V + cb - 1) / cb);
    hipLaunchKernelGGL(k_chansum, dim3(nb), dim3(64), 0, U->stream, gout, V, c.cout, cb, U->part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, U->stream, U->part, nb,
                       (int64_t)c.cout, c.gb);
}

void relu_mask(xh_unet_s *U, int64_t n, const bf16_t *act, bf16_t *g) {
    hipLaunchKernelGGL(k_relu_mask, dim3(grid1(n)), dim3(256), 0, U->stream, n, act, g);
}

// 0: inference materialises the upsampled decoder inputs as training does
int g_fuse_up = 1;

// infer: the upsampled decoder inputs are not kept (no backward follows): the decoder's
// first convolution reads the coarse tensor through the fused upsampling
void forward(xh_unet_s *U, Work &W, bool infer = false) {
    const int L = U->L;
    hipStream_t st = U->stream;
    for (int l = 0; l < L; ++l) {
        const Dims &g = W.g[l];
        conv_fwd(U, U->enc[2 * l], g, W.in[l], nullptr, true, W.a[l]);
        conv_fwd(U, U->enc[2 * l + 1], g, W.a[l], nullptr, true, W.s[l]);
        if (l + 1 < L) {
            const int64_t n = W.g[l + 1].vox() * U->ch[l];
            hipLaunchKernelGGL(k_pool, dim3(grid1(n)), dim3(256), 0, st, g, U->ch[l], W.s[l], W.in[l + 1]);
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        const Dims &g = W.g[l];
        const bf16_t *hup = l + 1 == L - 1 ? W.s[L - 1] : W.hh[l + 1];
        if (infer && g_fuse_up) {
            conv_fwd(U, U->dec[2 * l], g, hup, W.s[l], true, W.dd[l], true);
        } else {
            const int64_t n = g.vox() * U->ch[l + 1];
            hipLaunchKernelGGL(k_up, dim3(grid1(n)), dim3(256), 0, st, g, U->ch[l + 1], hup,
                               W.u[l]);
            conv_fwd(U, U->dec[2 * l], g, W.u[l], W.s[l], true, W.dd[l]);
        }
        conv_fwd(U, U->dec[2 * l + 1], g, W.dd[l], nullptr, true, W.hh[l]);
    }
    if (infer)
        return; // the head runs fused with the output scaling (k_head_out)
    const bf16_t *h0 = L == 1 ? W.s[0] : W.hh[0];
    const int64_t V0 = W.g[0].vox();
    hipLaunchKernelGGL(k_head, dim3(grid1(V0)), dim3(256), 0, st, V0, U->hc, h0, U->hw, U->hb, W.y);
}

// loss into U->dloss (device), gy
void loss(xh_unet_s *U, Work &W) {
    const int64_t V0 = W.g[0].vox();
    const int64_t cb = (V0 + 1023) / 1024;
    const int nb = (int)((V0 + cb - 1) / cb);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, U->stream, V0, W.y, W.tgt,
                       (float)(1.0 / (double)V0), W.gy, cb, U->part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, U->stream, U->part, nb,
                       (int64_t)1, U->dloss);
}

void backward(xh_unet_s *U, Work &W) {
    const int L = U->L;
    hipStream_t st = U->stream;
    const int64_t V0 = W.g[0].vox();
    // head
    {
        const bf16_t *h0 = L == 1 ? W.s[0] : W.hh[0];
        bf16_t *gh0 = L == 1 ? W.gs[0] : W.ghh[0];
        const int64_t cb = (V0 + 1023) / 1024;
        const int nb = (int)((V0 + cb - 1) / cb);
        hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, st, V0, U->hc, h0, U->hw, W.gy,
                           gh0, cb, U->part);
        // part[chunk][0..C]: gw then gb
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, st, U->part, nb,
                           (int64_t)(U->hc + 1), U->hgw); // hgw has hc+1 slots: w then b
    }
    // decoder, top level first
    for (int l = 0; l + 1 < L; ++l) {
        const Dims &g = W.g[l];
        const int64_t V = g.vox();
        const int cl = U->ch[l];
        relu_mask(U, V * cl, W.hh[l], W.ghh[l]);
        conv_wgrad(U, U->dec[2 * l + 1], g, W.dd[l], W.ghh[l]);
        conv_data(U, U->dec[2 * l + 1], g, W.ghh[l], W.gdd[l]);
        relu_mask(U, V * cl, W.dd[l], W.gdd[l]);
        conv_wgrad(U, U->dec[2 * l], g, W.u[l], W.gdd[l]);
        // skip path: the masked gradient also flows into the encoder output s_l
        (void)hipMemcpyAsync(W.gs[l], W.gdd[l], V * cl * sizeof(bf16_t), hipMemcpyDeviceToDevice, st);
        conv_data(U, U->dec[2 * l], g, W.gdd[l], W.gu[l]);
        bf16_t *gup = l + 1 == L - 1 ? W.gs[L - 1] : W.ghh[l + 1];
        const int64_t nc = W.g[l + 1].vox() * U->ch[l + 1];
        hipLaunchKernelGGL(k_up_bwd, dim3(grid1(nc)), dim3(256), 0, st, g, U->ch[l + 1], W.gu[l], gup);
    }
    // encoder, bottom level first; gs[l] holds the skip gradient (or the whole gradient
    // at the bottom), the pooled path adds pool_bwd(gin[l+1])
    for (int l = L - 1; l >= 0; --l) {
        const Dims &g = W.g[l];
        const int64_t V = g.vox();
        const int cl = U->ch[l];
        if (l + 1 < L) {
            hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(V * cl)), dim3(256), 0, st, g, cl, W.gin[l + 1], W.tmp[l]);
            hipLaunchKernelGGL(k_add, dim3(grid1(V * cl)), dim3(256), 0, st, V * cl, W.gs[l], W.tmp[l]);
        }
        relu_mask(U, V * cl, W.s[l], W.gs[l]);
        conv_wgrad(U, U->enc[2 * l + 1], g, W.a[l], W.gs[l]);
        conv_data(U, U->enc[2 * l + 1], g, W.gs[l], W.ga[l]);
        relu_mask(U, V * cl, W.a[l], W.ga[l]);
        conv_wgrad(U, U->enc[2 * l], g, W.in[l], W.ga[l]);
        if (l > 0)
            conv_data(U, U->enc[2 * l], g, W.ga[l], W.gin[l]);
    }
}

void adam(xh_unet_s *U, float lr) {
    U->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)U->adam_t), c2 = 1.f - std::pow(b2, (float)U->adam_t);
    auto step = [&](int64_t n, float *p, const float *g, float *m, float *v) {
        hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, U->stream, n, p, g, m, v, lr,
                           b1, b2, c1, c2, eps);
    };
    for (std::vector<Conv> *cs : {&U->enc, &U->dec})
        for (Conv &c : *cs) {
            step(c.nw, c.w, c.gw, c.mw, c.vw);
            step(c.cout, c.b, c.gb, c.mb, c.vb);
        }
    step(U->hc, U->hw, U->hgw, U->hmw, U->hvw);
    step(1, U->hb, U->hgw + U->hc, U->hmb, U->hvb);
    prep_all(U);
}

// flat parameter (or gradient) copy between host and device
int copy_params(xh_unet_s *U, float *host, const float *chost, bool grads, bool to_host) {
    size_t off = 0;
    auto mv = [&](float *dev, int64_t n) -> int {
        if (to_host)
            UN_HIP(hipMemcpy(host + off, dev, n * 4, hipMemcpyDeviceToHost));
        else
            UN_HIP(hipMemcpy(dev, chost + off, n * 4, hipMemcpyHostToDevice));
        off += n;
        return 0;
    };
    int rc;
    for (std::vector<Conv> *cs : {&U->enc, &U->dec})
        for (Conv &c : *cs)
            if ((rc = mv(grads ? c.gw : c.w, c.nw)) || (rc = mv(grads ? c.gb : c.b, c.cout)))
                return rc;
    if ((rc = mv(grads ? U->hgw : U->hw, U->hc)) || (rc = mv(grads ? U->hgw + U->hc : U->hb, 1)))
        return rc;
    return 0;
}

void free_unet(xh_unet_s *U) {
    if (U->stream)
        (void)hipStreamSynchronize(U->stream);
    for (auto &w : U->work)
        for (void *p : w->owned)
            (void)hipFree(p);
    for (void *p : U->owned)
        (void)hipFree(p);
    if (U->stream)
        (void)hipStreamDestroy(U->stream);
    delete U;
}

} // namespace

void set_fuse_up(int on) { g_fuse_up = on ? 1 : 0; }
void set_conv_tile(int on) { g_conv_tile = on ? 1 : 0; }

} // namespace unet
} // namespace xh

using namespace xh;
using namespace xh::unet;

extern "C" int xh_unet_create(int levels, const int32_t *channels, uint64_t seed, xh_unet *out) {
    int rc = runtime_init();
    if (rc)
        return rc;
    if (levels < 1 || levels > 6 || !channels || !out)
        return report_error(XH_EINVAL, "xh_unet_create: 1..6 levels and channels needed");
    for (int l = 0; l < levels; ++l)
        if (channels[l] < 16 || channels[l] > 64 || channels[l] % 16)
            return report_error(XH_EINVAL, "xh_unet_create: channels must be 16, 32, 48 or 64");
    auto *U = new xh_unet_s();
    U->L = levels;
    U->ch.assign(channels, channels + levels);
    uint64_t draw = 0;
    hipError_t e = hipStreamCreate(&U->stream);
    if (e != hipSuccess) {
        delete U;
        return report_error(XH_EHIP, "xh_unet_create: stream");
    }
    auto bail = [&](int code) {
        free_unet(U);
        return code;
    };
    for (int l = 0; l < levels; ++l) {
        Conv a, b;
        if ((rc = make_conv(U, l == 0 ? 1 : U->ch[l - 1], U->ch[l], seed, draw, a)) ||
            (rc = make_conv(U, U->ch[l], U->ch[l], seed, draw, b)))
            return bail(rc);
        U->enc.push_back(a);
        U->enc.push_back(b);
    }
    for (int l = 0; l + 1 < levels; ++l) {
        Conv u, d;
        if ((rc = make_conv(U, U->ch[l + 1], U->ch[l], seed, draw, u)) ||
            (rc = make_conv(U, U->ch[l], U->ch[l], seed, draw, d)))
            return bail(rc);
        U->dec.push_back(u);
        U->dec.push_back(d);
    }
    U->hc = U->ch[0];
    for (float **q : {&U->hw, &U->hmw, &U->hvw})
        if ((rc = dalloc_t(U, q, U->hc, U->owned)))
            return bail(rc);
    if ((rc = dalloc_t(U, &U->hgw, U->hc + 1, U->owned)))
        return bail(rc);
    for (float **q : {&U->hb, &U->hmb, &U->hvb, &U->hgb})
        if ((rc = dalloc_t(U, q, 1, U->owned)))
            return bail(rc);
    std::vector<float> hw(U->hc);
    for (int c = 0; c < U->hc; ++c)
        hw[c] = (float)(std::sqrt(1.0 / U->hc) * normal(seed, draw++));
    if (hipMemcpy(U->hw, hw.data(), U->hc * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(report_error(XH_EHIP, "xh_unet_create: upload"));
    int64_t np = 0;
    for (auto *cs : {&U->enc, &U->dec})
        for (auto &c : *cs)
            np += c.nw + c.cout;
    U->nparams = np + U->hc + 1;
    U->part_len = (size_t)kMaxChunks * 27 * 64 * 64 + 4096;
    if ((rc = dalloc_t(U, &U->part, U->part_len, U->owned)) || (rc = dalloc_t(U, &U->scale, 64, U->owned)) ||
        (rc = dalloc_t(U, &U->dloss, 1, U->owned)) ||
        (rc = dalloc_t(U, &U->dpart, 1024 * 64, U->owned)))
        return bail(rc);
    prep_all(U);
    if (hipStreamSynchronize(U->stream) != hipSuccess)
        return bail(report_error(XH_EHIP, "xh_unet_create: init"));
    *out = U;
    return 0;
}

extern "C" int xh_unet_destroy(xh_unet U) {
    if (U)
        free_unet(U);
    return 0;
}

extern "C" int xh_unet_nparams(xh_unet U, int64_t *n) {
    if (!U || !n)
        return report_error(XH_EINVAL, "xh_unet_nparams: bad arguments");
    *n = U->nparams;
    return 0;
}

extern "C" int xh_unet_get_params(xh_unet U, float *p) {
    if (!U || !p)
        return report_error(XH_EINVAL, "xh_unet_get_params: bad arguments");
    return copy_params(U, p, nullptr, false, true);
}

extern "C" int xh_unet_set_params(xh_unet U, const float *p) {
    if (!U || !p)
        return report_error(XH_EINVAL, "xh_unet_set_params: bad arguments");
    int rc = copy_params(U, nullptr, p, false, false);
    if (rc)
        return rc;
    prep_all(U);
    UN_HIP(hipStreamSynchronize(U->stream));
    return 0;
}

extern "C" int xh_unet_get_output_scale(xh_unet U, float *o) {
    if (!U || !o)
        return report_error(XH_EINVAL, "xh_unet_get_output_scale: bad arguments");
    *o = U->oscale;
    return 0;
}

extern "C" int xh_unet_set_output_scale(xh_unet U, float o) {
    if (!U || !(o > 0.f))
        return report_error(XH_EINVAL, "xh_unet_set_output_scale: scale must be > 0");
    U->oscale = o;
    return 0;
}

extern "C" int xh_unet_loss_grad(xh_unet U, int n, int d, int h, int w, const float *in,
                                 const float *tgt, float *loss_out, float *grads, float *y) {
    if (!U || n < 1 || !in || !tgt || !level_dims_ok(U, d, h, w))
        return report_error(XH_EINVAL, "xh_unet_loss_grad: bad arguments (dims must be "
                                       "multiples of 2^(levels-1))");
    Work *W = nullptr;
    int rc = get_work(U, n, d, h, w, true, &W);
    if (rc)
        return rc;
    const int64_t V0 = W->g[0].vox();
    UN_HIP(hipMemcpy(W->tgt, in, V0 * 4, hipMemcpyHostToDevice)); // staging
    hipLaunchKernelGGL(k_f2bf, dim3(grid1(V0)), dim3(256), 0, U->stream, V0, W->tgt, W->in[0]);
    UN_HIP(hipStreamSynchronize(U->stream));
    UN_HIP(hipMemcpy(W->tgt, tgt, V0 * 4, hipMemcpyHostToDevice));
    forward(U, *W);
    loss(U, *W);
    backward(U, *W);
    UN_HIP(hipGetLastError());
    UN_HIP(hipStreamSynchronize(U->stream));
    if (loss_out)
        UN_HIP(hipMemcpy(loss_out, U->dloss, 4, hipMemcpyDeviceToHost));
    if (y)
        UN_HIP(hipMemcpy(y, W->y, V0 * 4, hipMemcpyDeviceToHost));
    if (grads)
        return copy_params(U, grads, nullptr, true, true);
    return 0;
}

// per-column RMS of a device multivector [V][nv] into U->scale
static void column_rms(xh_unet_s *U, int64_t V, int nv, const double *b) {
    const int64_t cb = (V + 1023) / 1024;
    const int nb = (int)((V + cb - 1) / cb);
    hipLaunchKernelGGL(k_rms_part, dim3(nb), dim3(256), 0, U->stream, V, nv, b, cb, U->dpart);
    hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(256), 0, U->stream, nb, nv, V, U->dpart, U->scale);
}

extern "C" int xh_unet_train(xh_unet U, int nx, int ny, int nz, int nsets,
                             const double *const *db, const double *const *dx, int nv, int crop,
                             int batch, int steps, float lr, uint64_t seed, float *loss_hist) {
    if (!U || nsets < 1 || !db || !dx || nv < 1 || nv > 64 || batch < 1 || steps < 0 ||
        !level_dims_ok(U, crop, crop, crop) || crop > nx || crop > ny || crop > nz)
        return report_error(XH_EINVAL, "xh_unet_train: bad arguments (crop must be a multiple "
                                       "of 2^(levels-1) within the grid)");
    const int64_t V = (int64_t)nx * ny * nz;
    const int64_t S = (int64_t)nsets * nv;
    // normalised samples: input b/s (bf16) and target x/s (fp32), [sample][V]
    bf16_t *in_all = nullptr;
    float *tgt_all = nullptr;
    std::vector<void *> own;
    int rc;
    if ((rc = dalloc_t(U, &in_all, S * V, own)) || (rc = dalloc_t(U, &tgt_all, S * V, own))) {
        for (void *p : own)
            (void)hipFree(p);
        return rc;
    }
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(U->stream);
        for (void *p : own)
            (void)hipFree(p);
        return code;
    };
    for (int q = 0; q < nsets; ++q) {
        column_rms(U, V, nv, db[q]);
        hipLaunchKernelGGL(k_gather_in, dim3((unsigned)((V + 63) / 64)), dim3(256), 0, U->stream, V, nv,
                           db[q], U->scale, in_all + (size_t)q * nv * V, dx[q],
                           tgt_all + (size_t)q * nv * V);
    }
    // output scale: RMS of the normalised targets
    {
        std::vector<float> tg((size_t)std::min<int64_t>(S * V, 1 << 22));
        if (hipMemcpyAsync(tg.data(), tgt_all, tg.size() * 4, hipMemcpyDeviceToHost, U->stream) != hipSuccess ||
            hipStreamSynchronize(U->stream) != hipSuccess)
            return cleanup(report_error(XH_EHIP, "xh_unet_train: targets"));
        double ss = 0.0;
        for (float v : tg)
            ss += (double)v * v;
        const double r = std::sqrt(ss / (double)tg.size());
        U->oscale = r > 0 ? (float)r : 1.f;
        // targets in units of the output scale
        hipLaunchKernelGGL(k_scale_f32, dim3(grid1(S * V)), dim3(256), 0, U->stream, S * V,
                           tgt_all, 1.f / U->oscale);
    }
    Work *W = nullptr;
    if ((rc = get_work(U, batch, crop, crop, crop, true, &W)))
        return cleanup(rc);
    int *d_org = nullptr;
    float *d_loss = nullptr;
    if ((rc = dalloc_t(U, &d_org, 4 * batch, own)) ||
        (rc = dalloc_t(U, &d_loss, std::max(1, steps), own)))
        return cleanup(rc);
    std::vector<int> org(4 * batch);
    const Dims gfull{(int)S, nz, ny, nx}, gcrop{batch, crop, crop, crop};
    const int64_t Vc = gcrop.vox();
    uint64_t draw = 0;
    for (int it = 0; it < steps; ++it) {
        for (int b = 0; b < batch; ++b) {
            org[4 * b] = (int)(unif(seed, draw++) * S);
            org[4 * b + 1] = (int)(unif(seed, draw++) * (nz - crop + 1));
            org[4 * b + 2] = (int)(unif(seed, draw++) * (ny - crop + 1));
            org[4 * b + 3] = (int)(unif(seed, draw++) * (nx - crop + 1));
        }
        // the previous step's kernels read d_org: finish them before overwriting it
        if (hipStreamSynchronize(U->stream) != hipSuccess ||
            hipMemcpy(d_org, org.data(), org.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(report_error(XH_EHIP, "xh_unet_train: origins"));
        hipLaunchKernelGGL(k_crop, dim3(grid1(Vc)), dim3(256), 0, U->stream, gfull, gcrop, d_org,
                           in_all, tgt_all, W->in[0], W->tgt);
        forward(U, *W);
        loss(U, *W);
        backward(U, *W);
        adam(U, lr);
        if (hipMemcpyAsync(d_loss + it, U->dloss, 4, hipMemcpyDeviceToDevice, U->stream) != hipSuccess ||
            hipGetLastError() != hipSuccess)
            return cleanup(report_error(XH_EHIP, "xh_unet_train: launch"));
    }
    if (hipStreamSynchronize(U->stream) != hipSuccess)
        return cleanup(report_error(XH_EHIP, "xh_unet_train: sync"));
    if (loss_hist && steps > 0 &&
        hipMemcpy(loss_hist, d_loss, steps * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return cleanup(report_error(XH_EHIP, "xh_unet_train: loss history"));
    return cleanup(0);
}

extern "C" int xh_unet_infer(xh_unet U, int nx, int ny, int nz, int nv, const double *db,
                             double *dx0) {
    if (!U || nv < 1 || nv > 64 || !db || !dx0 || !level_dims_ok(U, nz, ny, nx))
        return report_error(XH_EINVAL, "xh_unet_infer: bad arguments (grid dims must be "
                                       "multiples of 2^(levels-1))");
    const int64_t V = (int64_t)nx * ny * nz;
    Work *W = nullptr;
    int rc = get_work(U, nv, nz, ny, nx, false, &W);
    if (rc)
        return rc;
    column_rms(U, V, nv, db);
    hipLaunchKernelGGL(k_gather_in, dim3((unsigned)((V + 63) / 64)), dim3(256), 0, U->stream, V, nv, db,
                       U->scale, W->in[0], nullptr, nullptr);
    forward(U, *W, true);
    // x0 = s * o * (head of the top decoder level)
    const bf16_t *h0 = U->L == 1 ? W->s[0] : W->hh[0];
    if (256 % nv == 0 && U->hc % 8 == 0)
        hipLaunchKernelGGL(k_head_out_t, dim3((unsigned)((V * nv + 255) / 256)), dim3(256), 0,
                           U->stream, V, nv, U->hc, h0, U->hw, U->hb, U->scale, U->oscale, dx0);
    else
        hipLaunchKernelGGL(k_head_out, dim3(grid1(V * nv)), dim3(256), 0, U->stream, V, nv,
                           U->hc, h0, U->hw, U->hb, U->scale, U->oscale, dx0);
    UN_HIP(hipGetLastError());
    UN_HIP(hipStreamSynchronize(U->stream));
    return 0;
}

extern "C" int xh_unet_mfma_selftest(const float *a, const float *b, float *d) {
    int rc = runtime_init();
    if (rc)
        return rc;
    if (!a || !b || !d)
        return report_error(XH_EINVAL, "xh_unet_mfma_selftest: bad arguments");
    float *da = nullptr, *dbv = nullptr, *dd = nullptr;
    UN_HIP(hipMalloc(&da, 16 * 32 * 4));
    UN_HIP(hipMalloc(&dbv, 32 * 16 * 4));
    UN_HIP(hipMalloc(&dd, 16 * 16 * 4));
    hipError_t e = hipMemcpy(da, a, 16 * 32 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(dbv, b, 32 * 16 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, da, dbv, dd);
        e = hipDeviceSynchronize();
    }
    if (e == hipSuccess)
        e = hipMemcpy(d, dd, 16 * 16 * 4, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(dbv);
    (void)hipFree(dd);
    if (e != hipSuccess)
        return report_error(XH_EHIP, std::string("xh_unet_mfma_selftest: ") + hipGetErrorString(e));
    return 0;
}
