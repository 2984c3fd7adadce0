// denoise.hip -- BF16 3D U-Net fluence denoiser: MFMA implicit-GEMM
// convolution kernels (forward, data gradient, weight gradient), the
// pointwise layers, the loss, Adam, and the C-ABI of tam_denoise.h.
//
// Layout: activations are channels-last BF16, [z][y][x][c] (x fastest over
// voxels, c fastest overall); a 3x3x3 convolution with zero padding is the
// GEMM  Y[m][co] = sum_k X_im2col[m][k] W[k][co],  k = tap * Cin + ci,
// tap = 9 (dz+1) + 3 (dy+1) + (dx+1).  One v_mfma_f32_16x16x32_bf16 takes a
// 16-voxel x 32-k slab of X (lane l: voxel l%16, k 8 (l/16) .. +7, one
// 16-byte load of 8 consecutive channels) and a 32 x 16 slab of W (packed
// on the host side of each update into the lane order, one 16-byte load),
// and accumulates a 16 x 16 FP32 tile (lane l: voxels 4 (l/16) + j, output
// channel l%16).  Cin is a multiple of 16, so 8 consecutive k never cross
// a tap.
//
// The data gradient is the same kernel with the weights flipped in space and
// transposed (W'[26 - tap][co][ci]).  The weight gradient needs k = voxels:
// each wave stages 32 consecutive voxels of the layer input (shifted by one
// tap, zero outside) and of the output gradient in LDS, transposed to
// [channel][voxel], and reads 8 consecutive voxels of one channel per lane.
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "tam_internal.h"
#include "../../include/tam_denoise.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

struct ConvArgs {
    const uint16_t* x;      // input [nvox][cin]; with x2: the coarse grid [nvox / 8][c1]
    const uint16_t* wp;     // packed weights [kb][nt][64][8]
    const float* bias;      // [cout] or nullptr
    const uint16_t* mask;   // [nvox][cout] or nullptr: output *= (mask > 0)
    uint16_t* y;            // output [nvox][cout]
    int D, H, W, cin, cout, kb, relu;
    // input = cat(upsample2(x), x2) when x2 != nullptr: channels < c1 from the
    // coarse grid (nearest), the other cin - c1 from x2 [nvox][cin - c1]
    const uint16_t* x2;
    int c1;
};

// 8 channels (c0 .. c0 + 7, c0 a multiple of 8) of fine voxel (xx, yy, zz):
// from x, or with x2 from cat(upsample2(x), x2)
__device__ __forceinline__ const uint16_t* cat_src(const uint16_t* x, const uint16_t* x2, int c1, int cin, int D,
                                                   int H, int W, int xx, int yy, int zz, int c0)
{
    if (!x2) return x + (((long)zz * H + yy) * W + xx) * cin + c0;
    if (c0 < c1) return x + (((long)(zz >> 1) * (H >> 1) + (yy >> 1)) * (W >> 1) + (xx >> 1)) * c1 + c0;
    return x2 + (((long)zz * H + yy) * W + xx) * (cin - c1) + (c0 - c1);
}

// Raw buffer descriptor over `bytes` bytes from p (at most 2^31 - 1): loads
// at offsets past it return zero.  0x00020000: gfx9 buffer word 3 (32-bit
// data format, untyped access).
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned v4u __attribute__((ext_vector_type(4)));
__device__ __forceinline__ Rsrc rsrc_n(const void* p, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(bytes < 0x7fffffffL ? bytes : 0x7fffffffL),
                                             0x00020000);
}
__device__ __forceinline__ bf16x8 bload8(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

// Forward convolution (and, with flipped weights, the data gradient).
// A wave computes MT x 16 consecutive voxels of the flattened grid (each
// lane finds its voxel's x, y, z) for all NT x 16 output channels; CIN is
// the input channel count.
// Addressing: the input descriptor starts HW + W + 1 voxels before the
// block's first voxel, so every tap of the block's voxels sits at a
// non-negative 32-bit offset; a tap outside the grid (zero padding) gets an
// offset past the descriptor's end and loads zeros.  Each lane keeps a
// 27-bit mask of its voxel's valid taps.  With the k loop unrolled, a
// k-block's tap and channel for each of the four lane groups are constants,
// and their offsets scalar values.
template <int CIN, int NT, int MT>
__global__ __launch_bounds__(256) void conv_kernel(ConvArgs a)
{
    constexpr int KB = (27 * CIN + 31) / 32;
    constexpr unsigned OUT = 0x80000000u;          // past every descriptor's end
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const int W = a.W, H = a.H;
    const long HW = (long)H * W, nvox = HW * a.D;
    const long mblk = (long)blockIdx.x * (4 * 16 * MT);
    const long mbase = mblk + (long)(threadIdx.x >> 6) * (16 * MT);
    if (mbase >= nvox) return;
    const long back = HW + W + 1;
    const long vb = mblk > back ? mblk - back : 0;
    const Rsrc rx = rsrc_n(a.x + vb * CIN, (nvox - vb) * CIN * 2);
    const Rsrc rw = rsrc_n(a.wp, (long)KB * NT * 512 * 2);
    unsigned voff[MT], vmask[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const long m = mbase + t * 16 + (lane & 15);
        const int vx = (int)(m % W);
        const long r = m / W;
        const int vy = (int)(r % H), vz = (int)(r / H);
        voff[t] = (unsigned)((m - vb) * CIN * 2);
        const unsigned bx = (vx > 0 ? 1u : 0u) | 2u | (vx < W - 1 ? 4u : 0u);
        const unsigned by = (vy > 0 ? 1u : 0u) | 2u | (vy < H - 1 ? 4u : 0u);
        const unsigned bz = (vz > 0 ? 1u : 0u) | 2u | (vz < a.D - 1 ? 4u : 0u);
        unsigned vm = 0;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
                if ((bz >> dz) & (by >> dy) & 1u) vm |= bx << (9 * dz + 3 * dy);
        vmask[t] = vm;
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        // lane group gg holds k = 32 kb + 8 gg .. + 7: one tap, 8 channels
        int tapg[4], offg[4];
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) {
            const int k = 32 * kb + 8 * gg;
            const int tap = k / CIN, ci = k % CIN;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            tapg[gg] = tap;                           // 27 (past the last tap) never has its mask bit set
            offg[gg] = (int)(((dz * HW + (long)dy * W + dx) * CIN + ci) * 2);
        }
        const int tap = g == 0 ? tapg[0] : g == 1 ? tapg[1] : g == 2 ? tapg[2] : tapg[3];
        const int toff = g == 0 ? offg[0] : g == 1 ? offg[1] : g == 2 ? offg[2] : offg[3];
        bf16x8 af[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const bool valid = (vmask[t] >> tap) & 1u;
            af[t] = bload8(rx, valid ? voff[t] + (unsigned)toff : OUT, 0);
        }
        bf16x8 bfr[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) bfr[n] = bload8(rw, (unsigned)lane * 16u, (unsigned)((kb * NT + n) * 1024));
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int n = 0; n < NT; ++n)
                acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[t], bfr[n], acc[t][n], 0, 0, 0);
    }
    // epilogue: lane holds voxels mbase + 16 t + 4 (lane/16) + j, channel 16 n + lane%16
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int co = n * 16 + (lane & 15);
        const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long m = mbase + t * 16 + (lane >> 4) * 4 + j;
                float v = acc[t][n][j] + b;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask && !(bf2f(a.mask[m * a.cout + co]) > 0.f)) v = 0.f;
                a.y[m * a.cout + co] = f2bf(v);
            }
    }
}

// Convolutions with at most 32 output channels, their input staged in LDS.
// In conv_kernel each A fragment of these layers feeds one or two MFMAs, so
// the vector-memory path, not the matrix cores, sets the rate.  Here a block
// computes an output brick of 16 (x) x 8 (y) x 4 (z) voxels.  For each group
// of 16 input channels it loads the input brick with a one-voxel halo
// (18 x 10 x 6 voxels x 16 channels, zeros outside the grid) into LDS, and
// every tap's A fragment is then an LDS read at a constant offset, with no
// per-tap validity test.  Wave w computes the 8 x-rows of plane z0 + w.  The
// k loop runs channel group by channel group (27 taps x 16 channels each, two
// taps per k-block), so each lane fetches its B fragment from the packed
// weights at that (tap, channel)'s place in their tap-major order.
// Grid edges must be multiples of 16 (x), 8 (y) and 4 (z); CIN a multiple of
// 16; NT = output channels / 16.
template <int CIN, int NT>
__global__ __launch_bounds__(256) void conv_lds_kernel(ConvArgs a)
{
    constexpr int BX = 18, BY = 10, BZ = 6, CG = CIN / 16;
    constexpr int KB = (27 * CIN + 31) / 32;                  // k-blocks of the packed weights
    __shared__ __attribute__((aligned(16))) uint16_t sx[BZ * BY * BX * 16];
    const int W = a.W, H = a.H, D = a.D;
    const int gx = W / 16, gy = H / 8;
    const int bxi = blockIdx.x % gx, byi = (blockIdx.x / gx) % gy, bzi = blockIdx.x / (gx * gy);
    const int x0 = bxi * 16, y0 = byi * 8, z0 = bzi * 4;
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15, w = threadIdx.x >> 6;
    const Rsrc rw = rsrc_n(a.wp, (long)KB * NT * 512 * 2);
    f32x4 acc[8][NT];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int cg = 0; cg < CG; ++cg) {
        if (cg > 0) __syncthreads();                          // the previous group's reads are done
        // stage channels 16 cg .. 16 cg + 15 of the input brick: 16-byte chunks, two per voxel
        for (int q = threadIdx.x; q < BZ * BY * BX * 2; q += 256) {
            const int vq = q >> 1, half = q & 1;
            const int lx = vq % BX, ly = (vq / BX) % BY, lz = vq / (BX * BY);
            const int xx = x0 - 1 + lx, yy = y0 - 1 + ly, zz = z0 - 1 + lz;
            bf16x8 v = {};
            if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
                v = *(const bf16x8*)cat_src(a.x, a.x2, a.c1, CIN, D, H, W, xx, yy, zz, 16 * cg + half * 8);
            *(bf16x8*)(sx + vq * 16 + half * 8) = v;
        }
        __syncthreads();
        // k-block kb: lane group g takes tap 2 kb + g / 2, channels 16 cg +
        // 8 (g % 2) .. + 7.  Tap 27 (the last block's second half) is the
        // packed weights' zero padding (or past the descriptor's end, which
        // reads zeros): its A fragment repeats tap 26, a finite value times 0.
        auto load = [&](int kb, bf16x8* af, bf16x8* bfr) {
            const int tapr = 2 * kb + (g >> 1), ci0 = 8 * (g & 1);
            const int tap = tapr < 27 ? tapr : 26;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const int base = ((w + dz + 1) * BY + (dy + 1)) * BX + (r + dx + 1);
#pragma unroll
            for (int t = 0; t < 8; ++t) af[t] = *(const bf16x8*)(sx + (base + t * BX) * 16 + ci0);
            // this lane's k in the packed (tap-major) order: its k-block and lane group there
            const int ko = tapr * CIN + 16 * cg + ci0;
            const int kbo = ko >> 5, go = (ko >> 3) & 3;
#pragma unroll
            for (int n = 0; n < NT; ++n)
                bfr[n] = bload8(rw, (unsigned)((go * 16 + r) * 16), (unsigned)((kbo * NT + n) * 1024));
        };
        auto mma = [&](const bf16x8* af, const bf16x8* bfr) {
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[t], bfr[n], acc[t][n], 0, 0, 0);
        };
        // two fragment sets in turn: the next k-block's reads are in flight
        // while the current one's MFMAs run
        bf16x8 afa[8], bfa[NT], afb[8], bfb[NT];
        load(0, afa, bfa);
        if constexpr (CG == 1 && NT == 1) {
            // unrolled: the accumulators stay in place (in a loop the compiler
            // rotates them through one register quad every iteration)
#pragma unroll
            for (int kb = 0; kb < 14; kb += 2) {
                load(kb + 1, afb, bfb);
                mma(afa, bfa);
                if (kb + 2 < 14) load(kb + 2, afa, bfa);
                mma(afb, bfb);
            }
        } else {
            // rolled: unrolled, the wider kernels grow past 160 VGPRs, and the
            // channel-group loop copies the accumulators anyway
#pragma unroll 1
            for (int kb = 0; kb < 14; kb += 2) {
                load(kb + 1, afb, bfb);
                mma(afa, bfa);
                if (kb + 2 < 14) load(kb + 2, afa, bfa);
                mma(afb, bfb);
            }
        }
    }
    // lane holds voxels x0 + 4 (lane / 16) + j of row (y0 + t, z0 + w), channel 16 n + lane % 16
    constexpr int COUT = NT * 16;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int co = n * 16 + (lane & 15);
        const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long m = ((long)(z0 + w) * H + (y0 + t)) * W + x0 + (lane >> 4) * 4 + j;
                float v = acc[t][n][j] + bb;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask && !(bf2f(a.mask[m * COUT + co]) > 0.f)) v = 0.f;
                a.y[m * COUT + co] = f2bf(v);
            }
    }
}

struct WgradArgs {
    const uint16_t* x;      // layer input [nvox][cin]
    const uint16_t* dy;     // output gradient [nvox][cout]
    float* gw;              // [27][cin][cout], accumulated with atomics
    int D, H, W, cin, cout;
    long blocks;            // 32-voxel segments of the flattened grid
    long per_wg;            // segments per workgroup (a multiple of 4)
};

// Weight gradient for one tap (blockIdx.y) over a range of 32-voxel
// segments (blockIdx.x): dW^T[co][ci] += sum_m dy[m][co] x[m + tap][ci].
template <int CIT, int COT>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a)
{
    constexpr int CI = CIT * 16, CO = COT * 16;
    constexpr int PAD = 8;                                  // row pad (elements) against bank conflicts
    __shared__ __attribute__((aligned(16))) uint16_t sX[4][CI][32 + PAD];
    __shared__ __attribute__((aligned(16))) uint16_t sD[4][CO][32 + PAD];
    __shared__ float sR[CI * CO];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    f32x4 acc[COT][CIT];
#pragma unroll
    for (int i = 0; i < COT; ++i)
#pragma unroll
        for (int c = 0; c < CIT; ++c) acc[i][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    const long b0 = (long)blockIdx.x * a.per_wg;
    const long b1 = b0 + a.per_wg < a.blocks ? b0 + a.per_wg : a.blocks;
    for (long bb = b0; bb < b1; bb += 4) {
        const long blk = bb + wv;
        const bool have = blk < b1;
        if (have) {
            // segment blk: the 32 consecutive voxels 32 blk .. 32 blk + 31 of
            // the flattened grid (each finds its own x, y, z below)
            const long mrow = blk * 32;
            // dy: 32 voxels x CO channels, 8 channels per lane-load
            for (int q = lane; q < 32 * CO / 8; q += 64) {
                const int v = q / (CO / 8), c8 = (q % (CO / 8)) * 8;
                const uint4 d = *(const uint4*)(a.dy + (mrow + v) * a.cout + c8);
                const uint16_t* e = (const uint16_t*)&d;
#pragma unroll
                for (int j = 0; j < 8; ++j) sD[wv][c8 + j][v] = e[j];
            }
            // x shifted by the tap, zero outside the grid
            for (int q = lane; q < 32 * CI / 8; q += 64) {
                const int v = q / (CI / 8), c8 = (q % (CI / 8)) * 8;
                const long m = mrow + v;
                const int x = (int)(m % a.W);
                const long r = m / a.W;
                const int sx = x + dx, sy = (int)(r % a.H) + dy, sz = (int)(r / a.H) + dz;
                uint4 d = make_uint4(0u, 0u, 0u, 0u);
                if (sx >= 0 && sx < a.W && sy >= 0 && sy < a.H && sz >= 0 && sz < a.D)
                    d = *(const uint4*)(a.x + (((long)sz * a.H + sy) * a.W + sx) * a.cin + c8);
                const uint16_t* e = (const uint16_t*)&d;
#pragma unroll
                for (int j = 0; j < 8; ++j) sX[wv][c8 + j][v] = e[j];
            }
        }
        __syncthreads();
        if (have) {
            const int k8 = (lane >> 4) * 8, r16 = lane & 15;
            bf16x8 bx[CIT];
#pragma unroll
            for (int c = 0; c < CIT; ++c) bx[c] = load8(&sX[wv][c * 16 + r16][k8]);
#pragma unroll
            for (int i = 0; i < COT; ++i) {
                const bf16x8 ad = load8(&sD[wv][i * 16 + r16][k8]);
#pragma unroll
                for (int c = 0; c < CIT; ++c)
                    acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad, bx[c], acc[i][c], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // reduce the four waves in LDS, then one atomic per weight
    for (int q = threadIdx.x; q < CI * CO; q += 256) sR[q] = 0.f;
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int i = 0; i < COT; ++i)
#pragma unroll
                for (int c = 0; c < CIT; ++c)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int co = i * 16 + (lane >> 4) * 4 + j, ci = c * 16 + (lane & 15);
                        sR[ci * CO + co] += acc[i][c][j];
                    }
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < CI * CO; q += 256)
        if (sR[q] != 0.f) atomicAdd(&a.gw[(long)tap * CI * CO + q], sR[q]);
}

// Weight gradient of the three taps dx = -1, 0, 1 that share (dz, dy) = blockIdx.y over a range of 32-voxel row
// segments (rows of W % 32 == 0 voxels, so a segment never leaves its row).  Each wave stages its segment's dy and
// the 34 x voxels x0-1 .. x0+32 of the source row (z+dz, y+dy) in its own LDS slice, channel-major; the operand
// of tap dx is the aligned 8-voxel fragment shifted by one bf16 (v_alignbyte on dword pairs), so one staging
// serves three taps.  Segments whose source row lies outside the grid contribute nothing and are skipped.
template <int CIT, int COT>
__global__ __launch_bounds__(256) void wgrad3_kernel(WgradArgs a)
{
    constexpr int CI = CIT * 16, CO = COT * 16;
    constexpr int XR = 48;                                  // x row: element 8 + v holds voxel x0 + v (v = -1 .. 32)
    constexpr int DR = 40;                                  // dy row: 32 voxels + pad against bank conflicts
    __shared__ __attribute__((aligned(16))) uint16_t sX[4][CI][XR];
    __shared__ __attribute__((aligned(16))) uint16_t sD[4][CO][DR];
    __shared__ float sR[CI * CO];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int dz = (int)blockIdx.y / 3 - 1, dy = (int)blockIdx.y % 3 - 1;
    f32x4 acc[3][COT][CIT];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int i = 0; i < COT; ++i)
#pragma unroll
            for (int c = 0; c < CIT; ++c) acc[s][i][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    const long b0 = (long)blockIdx.x * a.per_wg;
    const long b1 = b0 + a.per_wg < a.blocks ? b0 + a.per_wg : a.blocks;
    const int k8 = (lane >> 4) * 8, r16 = lane & 15;
    for (long blk = b0 + wv; blk < b1; blk += 4) {
        const long m0 = blk * 32;
        const int x0 = (int)(m0 % a.W);
        const long r = m0 / a.W;
        const int sy = (int)(r % a.H) + dy, sz = (int)(r / a.H) + dz;
        if (sy < 0 || sy >= a.H || sz < 0 || sz >= a.D) continue;      // wave-uniform
        // staging by voxel pairs: each lane loads 8 channels of two neighbouring voxels and writes one dword
        // (both voxels) per channel
        for (int q = lane; q < 16 * (CO / 8); q += 64) {
            const int p = q / (CO / 8), c8 = (q % (CO / 8)) * 8;
            const uint4 d0 = *(const uint4*)(a.dy + (m0 + 2 * p) * a.cout + c8);
            const uint4 d1 = *(const uint4*)(a.dy + (m0 + 2 * p + 1) * a.cout + c8);
            const uint16_t* e0 = (const uint16_t*)&d0;
            const uint16_t* e1 = (const uint16_t*)&d1;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                *(uint32_t*)&sD[wv][c8 + j][2 * p] = (uint32_t)e0[j] | ((uint32_t)e1[j] << 16);
        }
        const long srow = ((long)sz * a.H + sy) * a.W;
        // pairs p = 0 .. 17: voxels x0 + 2p - 2 and x0 + 2p - 1 at elements 2p + 6, 2p + 7 (elements 7 .. 40 are
        // the ones read; 6 and 41 come along)
        for (int q = lane; q < 18 * (CI / 8); q += 64) {
            const int p = q / (CI / 8), c8 = (q % (CI / 8)) * 8;
            const int sx = x0 + 2 * p - 2;
            uint4 d0 = make_uint4(0u, 0u, 0u, 0u), d1 = make_uint4(0u, 0u, 0u, 0u);
            if (sx >= 0 && sx < a.W) d0 = *(const uint4*)(a.x + (srow + sx) * a.cin + c8);
            if (sx + 1 >= 0 && sx + 1 < a.W) d1 = *(const uint4*)(a.x + (srow + sx + 1) * a.cin + c8);
            const uint16_t* e0 = (const uint16_t*)&d0;
            const uint16_t* e1 = (const uint16_t*)&d1;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                *(uint32_t*)&sX[wv][c8 + j][2 * p + 6] = (uint32_t)e0[j] | ((uint32_t)e1[j] << 16);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");         // the slice is written before it is read
        __builtin_amdgcn_wave_barrier();
        bf16x8 bx[3][CIT];
#pragma unroll
        for (int c = 0; c < CIT; ++c) {
            const uint16_t* row = &sX[wv][c * 16 + r16][0];
            const uint4 mid = *(const uint4*)(row + 8 + k8);                  // voxels k8 .. k8+7
            const unsigned prev = *(const unsigned*)(row + 6 + k8);          // high half: voxel k8-1
            const unsigned next = *(const unsigned*)(row + 16 + k8);         // low half: voxel k8+8
            uint4 lo, hi;
            lo.x = __builtin_amdgcn_alignbyte(mid.x, prev, 2);
            lo.y = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
            lo.z = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
            lo.w = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
            hi.x = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
            hi.y = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
            hi.z = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
            hi.w = __builtin_amdgcn_alignbyte(next, mid.w, 2);
            bx[0][c] = __builtin_bit_cast(bf16x8, lo);
            bx[1][c] = __builtin_bit_cast(bf16x8, mid);
            bx[2][c] = __builtin_bit_cast(bf16x8, hi);
        }
#pragma unroll
        for (int i = 0; i < COT; ++i) {
            const bf16x8 ad = load8(&sD[wv][i * 16 + r16][k8]);
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int c = 0; c < CIT; ++c)
                    acc[s][i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad, bx[s][c], acc[s][i][c], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");         // the reads finish before the next writes
        __builtin_amdgcn_wave_barrier();
    }
    // per tap: reduce the four waves in LDS, then one atomic per weight
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        for (int q = threadIdx.x; q < CI * CO; q += 256) sR[q] = 0.f;
        __syncthreads();
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
#pragma unroll
                for (int i = 0; i < COT; ++i)
#pragma unroll
                    for (int c = 0; c < CIT; ++c)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int co = i * 16 + (lane >> 4) * 4 + j, ci = c * 16 + (lane & 15);
                            sR[ci * CO + co] += acc[s][i][c][j];
                        }
            }
            __syncthreads();
        }
        const int tap = (int)blockIdx.y * 3 + s;
        for (int q = threadIdx.x; q < CI * CO; q += 256)
            if (sR[q] != 0.f) atomicAdd(&a.gw[(long)tap * CI * CO + q], sR[q]);
        __syncthreads();
    }
}

// Weight gradient by output bricks, all of a 16-channel input tile's taps at
// once: dW^T[co][ci] += sum_m dy[m][co] x[m + tap][ci].
//
// A block stages a 16 (x) x 8 (y) x 4 (z) brick of dy (COT x 16 channels) and
// the matching 18 x 10 x 6 brick of x (one 16-channel tile, zeros outside the
// grid), both transposed to [channel][voxel] in LDS, once per brick.  Wave w
// owns plane w of the brick: four 32-voxel k-chunks (row pairs), whose dy
// fragments stay in registers while it walks the taps.  Per chunk and (dz, dy)
// one aligned 8-voxel x read plus two dwords give the dx = -1, 0, +1 fragments
// (v_alignbyte), so one LDS read feeds 3 x COT MFMAs.  NDZ = 1: the block
// covers all 27 taps; NDZ = 3: the 9 taps of one dz (blockIdx.y % 3).
// blockIdx.y / NDZ is the input-channel tile.  Blocks loop over bricks and
// keep their sums in registers; at the end the four waves are summed in LDS
// and each block writes its partial [taps][16][CO] to `part`, which
// wgradb_reduce_kernel sums into gw.  The blocks of grid row 0 also sum dy
// over their bricks (the bias gradient, CO floats after the weights).
struct WgradbArgs {
    const uint16_t* x;      // layer input [nvox][cin]; with x2: the coarse grid [nvox / 8][c1]
    const uint16_t* dy;     // output gradient [nvox][cout]
    float* part;            // [gridDim.y][gridDim.x]: [taps][16][cout] weights, then [cout] bias
    int D, H, W, cin, cout;
    long bricks;            // (W / 16) (H / 8) (D / 4)
    const uint16_t* x2;     // input = cat(upsample2(x), x2) when not null (see ConvArgs)
    int c1;
};

template <int COT, int NDZ>
__global__ __launch_bounds__(256) void wgradb_kernel(WgradbArgs a)
{
    constexpr int CO = COT * 16, NT = 27 / NDZ;
    constexpr int XR = 24;                                  // x row: element 4 + v holds voxel x0 + v (v = -2 .. 17)
    constexpr int XC = 6 * 10 * XR + 4;                     // x channel stride (elements), padded against bank conflicts
    constexpr int DC = 4 * 8 * 16 + 8;                      // dy channel stride (elements), padded likewise
    constexpr int NW = NT * 16 * CO;                       // weights per partial
    constexpr int XBYTES = 16 * XC * 2, DBYTES = CO * DC * 2, RBYTES = (NW + CO) * 4;
    constexpr int LBYTES = (XBYTES + DBYTES > RBYTES ? XBYTES + DBYTES : RBYTES);
    __shared__ __attribute__((aligned(16))) unsigned char lds[LBYTES];
    uint16_t* sX = (uint16_t*)lds;
    uint16_t* sD = (uint16_t*)(lds + XBYTES);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4, h = g & 1, gr = g >> 1;
    const int ct = (int)blockIdx.y / NDZ;                   // input-channel tile
    const int dz0 = NDZ == 1 ? -1 : (int)blockIdx.y % 3 - 1;
    const int W = a.W, H = a.H, D = a.D;
    const int bx = W / 16, by = H / 8;
    f32x4 acc[NT][COT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < COT; ++i) acc[t][i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool bias = blockIdx.y == 0;
    float bsum[COT];
#pragma unroll
    for (int i = 0; i < COT; ++i) bsum[i] = 0.f;

    for (long b = blockIdx.x; b < a.bricks; b += gridDim.x) {
        const int x0 = (int)(b % bx) * 16, y0 = (int)((b / bx) % by) * 8, z0 = (int)(b / ((long)bx * by)) * 4;
        __syncthreads();                                    // the last brick's reads are done
        // x: voxel pairs (x0 + 2p, x0 + 2p + 1), p = -1 .. 8, of 10 rows x 6
        // planes, 16 channels as two 8-channel halves; each lane loads both
        // voxels' half and writes one dword (the pair) per channel.  Voxels
        // x0 - 2 and x0 + 17 come along; their halves are never used.
        for (int q = threadIdx.x; q < 10 * 10 * 6 * 2; q += 256) {
            const int vq = q >> 1, half = q & 1;
            const int pp = vq % 10, ly = (vq / 10) % 10, lz = vq / 100;
            const int xx = x0 + 2 * pp - 2, yy = y0 - 1 + ly, zz = z0 - 1 + lz;
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < H && zz >= 0 && zz < D) {
                const int c0 = ct * 16 + half * 8;
                if (xx >= 0 && xx < W)
                    v0 = *(const uint4*)cat_src(a.x, a.x2, a.c1, a.cin, D, H, W, xx, yy, zz, c0);
                if (xx + 1 >= 0 && xx + 1 < W)
                    v1 = *(const uint4*)cat_src(a.x, a.x2, a.c1, a.cin, D, H, W, xx + 1, yy, zz, c0);
            }
            const uint16_t* e0 = (const uint16_t*)&v0;
            const uint16_t* e1 = (const uint16_t*)&v1;
            uint16_t* dst = sX + (half * 8) * XC + (lz * 10 + ly) * XR + 2 + 2 * pp;
#pragma unroll
            for (int j = 0; j < 8; ++j) *(uint32_t*)(dst + j * XC) = (uint32_t)e0[j] | ((uint32_t)e1[j] << 16);
        }
        // dy: voxel pairs of 16 x 8 x 4 voxels, CO channels in 8-channel groups
        for (int q = threadIdx.x; q < 256 * (CO / 8); q += 256) {
            const int vp = q / (CO / 8), c8 = (q % (CO / 8)) * 8;
            const int vq = 2 * vp;
            const int lx = vq & 15, ly = (vq >> 4) & 7, lz = vq >> 7;
            const uint16_t* src = a.dy + (((long)(z0 + lz) * H + (y0 + ly)) * W + (x0 + lx)) * a.cout + c8;
            const uint4 v0 = *(const uint4*)src;
            const uint4 v1 = *(const uint4*)(src + a.cout);
            const uint16_t* e0 = (const uint16_t*)&v0;
            const uint16_t* e1 = (const uint16_t*)&v1;
            uint16_t* dst = sD + c8 * DC + vq;
#pragma unroll
            for (int j = 0; j < 8; ++j) *(uint32_t*)(dst + j * DC) = (uint32_t)e0[j] | ((uint32_t)e1[j] << 16);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            // dy^T fragment of chunk c: channel r16 of tile i, voxels 8h .. 8h + 7 of row 2c + gr, plane wv
            bf16x8 ad[COT];
#pragma unroll
            for (int i = 0; i < COT; ++i) ad[i] = load8(sD + (i * 16 + r16) * DC + (wv * 8 + 2 * c + gr) * 16 + 8 * h);
            if (bias) {
#pragma unroll
                for (int i = 0; i < COT; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) bsum[i] += (float)ad[i][j];
            }
            // all of the chunk's x reads first (one wait for the lot), then its MFMAs
            constexpr int NR = (NDZ == 1 ? 3 : 1) * 3;
            uint2 m0[NR], m1[NR];
            unsigned prev[NR], next[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                // x row of channel r16 at plane wv + dz, row 2c + gr + dy (halo coordinates +1)
                const int dz = dz0 + r / 3, ky = r % 3;
                const uint16_t* row = sX + r16 * XC + ((wv + dz + 1) * 10 + (2 * c + gr + ky)) * XR;
                m0[r] = *(const uint2*)(row + 4 + 8 * h);                       // voxels 8h .. 8h + 3
                m1[r] = *(const uint2*)(row + 8 + 8 * h);                       // voxels 8h + 4 .. 8h + 7
                prev[r] = *(const unsigned*)(row + 2 + 8 * h);                  // high half: voxel 8h - 1
                next[r] = *(const unsigned*)(row + 12 + 8 * h);                 // low half: voxel 8h + 8
            }
#pragma unroll
            for (int kz = 0; kz < (NDZ == 1 ? 3 : 1); ++kz) {
#pragma unroll
                for (int ky = 0; ky < 3; ++ky) {
                    const int r = kz * 3 + ky;
                    uint4 mid, lo, hi;
                    mid.x = m0[r].x; mid.y = m0[r].y; mid.z = m1[r].x; mid.w = m1[r].y;
                    lo.x = __builtin_amdgcn_alignbyte(mid.x, prev[r], 2);
                    lo.y = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
                    lo.z = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
                    lo.w = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
                    hi.x = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
                    hi.y = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
                    hi.z = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
                    hi.w = __builtin_amdgcn_alignbyte(next[r], mid.w, 2);
                    const bf16x8 bx3[3] = {__builtin_bit_cast(bf16x8, lo), __builtin_bit_cast(bf16x8, mid),
                                           __builtin_bit_cast(bf16x8, hi)};
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx)
#pragma unroll
                        for (int i = 0; i < COT; ++i) {
                            const int t = (kz * 3 + ky) * 3 + kx;
                            acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad[i], bx3[kx], acc[t][i], 0, 0, 0);
                        }
                }
            }
        }
    }
    // the four waves' sums: wave 0 stores, the others add in turn; then the
    // block's partial goes out (element (t, ci, co) at (t * 16 + ci) * CO + co)
    float* sR = (float*)lds;
    __syncthreads();                                        // the last brick's reads are done
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int i = 0; i < COT; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int q = (t * 16 + r16) * CO + i * 16 + 4 * g + j;
                        sR[q] = (w == 0 ? 0.f : sR[q]) + acc[t][i][j];
                    }
        }
        __syncthreads();
    }
    // bias: the four k-groups of a channel, then the waves in turn
#pragma unroll
    for (int i = 0; i < COT; ++i) {
        bsum[i] += __shfl_xor(bsum[i], 16, 64);
        bsum[i] += __shfl_xor(bsum[i], 32, 64);
    }
    for (int w = 0; w < 4; ++w) {
        if (wv == w && g == 0) {
#pragma unroll
            for (int i = 0; i < COT; ++i) sR[NW + i * 16 + r16] = (w == 0 ? 0.f : sR[NW + i * 16 + r16]) + bsum[i];
        }
        __syncthreads();
    }
    float* out = a.part + ((long)blockIdx.y * gridDim.x + blockIdx.x) * (NW + CO);
    for (int q = threadIdx.x; q < NW + CO; q += 256) out[q] = sR[q];
}

// gw[tap][ci][co] += the sum over the nb blocks of one blockIdx.y of their
// partials; ny rows of blocks, ndz dz groups per input-channel tile.  Block
// column blockIdx.y of this grid sums partials k = blockIdx.y, blockIdx.y +
// gridDim.y, ... and adds its share atomically, so a weight sums over enough
// blocks in parallel to hide the memory latency.
__global__ __launch_bounds__(256) void wgradb_reduce_kernel(const float* __restrict__ part, int nb, int ny, int ndz,
                                                            int cin, int cout, float* __restrict__ gw,
                                                            float* __restrict__ gb)
{
    const int nt = 27 / ndz;
    const long nw = (long)nt * 16 * cout, per = nw + cout;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= per * ny) return;
    const int yrow = (int)(q / per);
    const long e = q % per;
    if (e >= nw && yrow != 0) return;                       // only grid row 0 sums the bias
    const float* src = part + (long)yrow * nb * per + e;
    float s = 0.f;
    for (int k = (int)blockIdx.y; k < nb; k += (int)gridDim.y) s += src[(long)k * per];
    if (e >= nw) {
        atomicAdd(&gb[e - nw], s);
        return;
    }
    const int t = (int)(e / (16 * cout)), ci = (int)((e / cout) % 16), co = (int)(e % cout);
    const int ct = yrow / ndz, tap = (ndz == 1 ? 0 : (yrow % 3) * 9) + t;
    atomicAdd(&gw[((long)tap * cin + ct * 16 + ci) * cout + co], s);
}

// bias gradient: gb[c] += sum_m d[m][c], c a multiple of 8 (<= 2048); each lane reads 8 channels of a voxel
__global__ __launch_bounds__(256) void bgrad_kernel(const uint16_t* d, long nvox, int c, float* gb)
{
    __shared__ float s[256 * 8];
    const int g = c / 8, per = 256 / g;                     // 8-channel groups per voxel, voxels per pass
    const int cg = threadIdx.x % g, vi = threadIdx.x / g;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (vi < per)
        for (long m = (long)blockIdx.x * per + vi; m < nvox; m += (long)gridDim.x * per) {
            const uint4 u = *(const uint4*)(d + m * c + cg * 8);
            const uint16_t* e = (const uint16_t*)&u;
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += bf2f(e[j]);
        }
#pragma unroll
    for (int j = 0; j < 8; ++j) s[threadIdx.x * 8 + j] = acc[j];
    __syncthreads();
    for (int ch = threadIdx.x; ch < c; ch += 256) {
        float t = 0.f;
        for (int v = 0; v < per; ++v) t += s[(v * g + ch / 8) * 8 + ch % 8];
        atomicAdd(&gb[ch], t);
    }
}

// The pointwise layers move 8 consecutive BF16 channels (16 bytes) per
// thread: every channel count here is a multiple of 8, so a chunk never
// straddles two source tensors.
struct Bf8 {
    uint16_t h[8];
};
__device__ __forceinline__ Bf8 ld8(const uint16_t* p) { return __builtin_bit_cast(Bf8, *(const uint4*)p); }
__device__ __forceinline__ void st8(uint16_t* p, const Bf8& v) { *(uint4*)p = __builtin_bit_cast(uint4, v); }

// 2x2x2 average pooling; one thread per coarse voxel and 8-channel chunk
__global__ void pool_kernel(const uint16_t* x, int D, int H, int W, int C, uint16_t* y)
{
    const int C8 = C / 8, Wc = W / 2, Hc = H / 2;
    const long n = (long)(D / 2) * Hc * Wc * C8;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C8) * 8;
        long r = q / C8;
        const int xo = (int)(r % Wc); r /= Wc;
        const int yo = (int)(r % Hc);
        const int zo = (int)(r / Hc);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k) {
            const int xi = 2 * xo + (k & 1), yi = 2 * yo + ((k >> 1) & 1), zi = 2 * zo + (k >> 2);
            const Bf8 v = ld8(x + (((long)zi * H + yi) * W + xi) * C + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += bf2f(v.h[j]);
        }
        Bf8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o.h[j] = f2bf(0.125f * s[j]);
        st8(y + q * 8, o);
    }
}

// gradient of the upsampled part of a concatenation, summed onto the coarse
// grid and masked by the coarse ReLU output: d[mc][c] = (sum of the 8
// children's g[child][c]) * (a[mc][c] > 0), c < C1; one thread per coarse
// voxel and 8-channel chunk
__global__ void upcat_bwd_kernel(const uint16_t* g, int C, int C1, const uint16_t* a, int D, int H, int W,
                                 uint16_t* d)
{
    const int Dc = D / 2, Hc = H / 2, Wc = W / 2, C18 = C1 / 8;
    const long n = (long)Dc * Hc * Wc * C18;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C18) * 8;
        long r = q / C18;
        const int xc = (int)(r % Wc); r /= Wc;
        const int yc = (int)(r % Hc);
        const int zc = (int)(r / Hc);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k) {
            const int x = 2 * xc + (k & 1), yy = 2 * yc + ((k >> 1) & 1), z = 2 * zc + (k >> 2);
            const Bf8 v = ld8(g + (((long)z * H + yy) * W + x) * C + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += bf2f(v.h[j]);
        }
        const Bf8 av = ld8(a + q * 8);
        Bf8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o.h[j] = bf2f(av.h[j]) > 0.f ? f2bf(s[j]) : (uint16_t)0;
        st8(d + q * 8, o);
    }
}

// gradient of a ReLU output that feeds a skip connection and a pooling:
// d[m][c] = (g_cat[m][C1 + c] + g_pool[parent(m)][c] / 8) * (a[m][c] > 0);
// one thread per fine voxel and 8-channel chunk
__global__ void skip_pool_bwd_kernel(const uint16_t* gcat, int Ccat, int C1, const uint16_t* gpool,
                                     const uint16_t* a, int D, int H, int W, int C, uint16_t* d)
{
    const int C8 = C / 8;
    const long n = (long)D * H * W * C8;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C8) * 8;
        const long m = q / C8;
        const int x = (int)(m % W);
        const long r = m / W;
        const int yy = (int)(r % H), z = (int)(r / H);
        const long mp = ((long)(z / 2) * (H / 2) + yy / 2) * (W / 2) + x / 2;
        const Bf8 gc = ld8(gcat + m * Ccat + C1 + c), gp = ld8(gpool + mp * C + c), av = ld8(a + q * 8);
        Bf8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float sj = bf2f(gc.h[j]) + 0.125f * bf2f(gp.h[j]);
            o.h[j] = bf2f(av.h[j]) > 0.f ? f2bf(sj) : (uint16_t)0;
        }
        st8(d + q * 8, o);
    }
}

// features (4 x nvox FP32, or built from a context) -> 16-channel input
__global__ void pack_input_kernel(const float* f, long nvox, uint16_t* x0)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x) {
        uint16_t v[16];
        for (int c = 0; c < 16; ++c) v[c] = c < 4 ? f2bf(f[(long)c * nvox + m]) : (uint16_t)0;
        uint4* o = (uint4*)(x0 + m * 16);
        o[0] = *(uint4*)&v[0];
        o[1] = *(uint4*)&v[8];
    }
}

// dataset slot (4 BF16 features + BF16 target per voxel) -> 16-channel input
__global__ void pack_slot_kernel(const uint16_t* feat, long nvox, uint16_t* x0)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x) {
        uint16_t v[16];
        for (int c = 0; c < 16; ++c) v[c] = c < 4 ? feat[m * 4 + c] : (uint16_t)0;
        uint4* o = (uint4*)(x0 + m * 16);
        o[0] = *(uint4*)&v[0];
        o[1] = *(uint4*)&v[8];
    }
}

// features from a context: phi = L / (npk V), mu_a and mu_s from mu_t and
// the material's albedo
__global__ void ctx_features_kernel(const double* tally, const float* mut, const uint8_t* mat,
                                    const MatEntry* mtab, long nvox, double inv_nv, float* f)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x) {
        const float mu = mut[m], alb = mtab[mat[m]].albedo;
        const float mua = mu * (1.f - alb), mus = mu * alb;
        f[m] = (float)log1p(tally[m] * inv_nv);
        f[nvox + m] = log10f(fmaxf(mua, 1e-3f));
        f[2 * nvox + m] = log10f(fmaxf(mus, 1e-2f)) - 2.f;
        f[3 * nvox + m] = 5.f * (mtab[mat[m]].g - 0.8f);
    }
}

// the same features written straight into the 16-channel BF16 input (the
// forward-only path: no FP32 staging)
__global__ void ctx_input_kernel(const double* tally, const float* mut, const uint8_t* mat, const MatEntry* mtab,
                                 long nvox, double inv_nv, uint16_t* x0)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x) {
        const float mu = mut[m], alb = mtab[mat[m]].albedo;
        const float mua = mu * (1.f - alb), mus = mu * alb;
        uint16_t v[16];
        // round to FP32 first, as the staged path stores it: the empty asm keeps
        // the compiler from folding double -> float -> BF16 into one rounding
        float l = (float)log1p(tally[m] * inv_nv);
        asm volatile("" : "+v"(l));
        v[0] = f2bf(l);
        v[1] = f2bf(log10f(fmaxf(mua, 1e-3f)));
        v[2] = f2bf(log10f(fmaxf(mus, 1e-2f)) - 2.f);
        v[3] = f2bf(5.f * (mtab[mat[m]].g - 0.8f));
#pragma unroll
        for (int c = 4; c < 16; ++c) v[c] = 0;
        uint4* o = (uint4*)(x0 + m * 16);
        o[0] = *(uint4*)&v[0];
        o[1] = *(uint4*)&v[8];
    }
}

// features (4 x nvox FP32) + target log1p(phi_hi) from a context -> slot
__global__ void store_slot_kernel(const float* f, const double* tally_hi, double inv_nv_hi, long nvox,
                                  uint16_t* feat, float* tgt)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x) {
        for (int c = 0; c < 4; ++c) feat[m * 4 + c] = f2bf(f[(long)c * nvox + m]);
        tgt[m] = tally_hi ? (float)log1p(tally_hi[m] * inv_nv_hi) : f[4 * nvox + m];
    }
}

// head: out[m] = x0[m][0] + hb + sum_c hw[c] a[m][c]; optional L1 against
// an FP32 target with the loss summed into *loss
__global__ __launch_bounds__(256) void head_kernel(const uint16_t* x0, const uint16_t* a, const float* hw,
                                                   long nvox, float* out, const float* tgt, double* loss,
                                                   double* base)
{
    __shared__ double s[2][256];
    double l = 0.0, lb = 0.0;
    for (long m = (long)blockIdx.x * 256 + threadIdx.x; m < nvox; m += (long)gridDim.x * 256) {
        float v = bf2f(x0[m * 16]) + hw[16];
        const uint4 u0 = *(const uint4*)(a + m * 16), u1 = *(const uint4*)(a + m * 16 + 8);
        const uint16_t* e0 = (const uint16_t*)&u0;
        const uint16_t* e1 = (const uint16_t*)&u1;
#pragma unroll
        for (int c = 0; c < 8; ++c) v += hw[c] * bf2f(e0[c]);
#pragma unroll
        for (int c = 0; c < 8; ++c) v += hw[8 + c] * bf2f(e1[c]);
        out[m] = v;
        if (tgt) {
            const float t = tgt[m];
            l += fabsf(v - t);
            lb += fabsf(bf2f(x0[m * 16]) - t);
        }
    }
    if (!tgt) return;
    s[0][threadIdx.x] = l;
    s[1][threadIdx.x] = lb;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s[0][threadIdx.x] += s[0][threadIdx.x + o];
            s[1][threadIdx.x] += s[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(loss, s[0][0]);
        if (base) atomicAdd(base, s[1][0]);
    }
}

// head backward: g = sign(out - t) / nvox; d9[m][c] = g hw[c] (a > 0);
// ghw[c] += sum g a[m][c]; ghw[16] += sum g
__global__ __launch_bounds__(256) void head_bwd_kernel(const float* out, const float* tgt, const uint16_t* a,
                                                       const float* hw, long nvox, uint16_t* d9, float* ghw)
{
    __shared__ float s[17][256];
    float acc[17];
    for (int c = 0; c < 17; ++c) acc[c] = 0.f;
    const float inv = 1.f / (float)nvox;
    for (long m = (long)blockIdx.x * 256 + threadIdx.x; m < nvox; m += (long)gridDim.x * 256) {
        const float diff = out[m] - tgt[m];
        const float g = (diff > 0.f ? inv : (diff < 0.f ? -inv : 0.f));
        uint4 u[2] = {*(const uint4*)(a + m * 16), *(const uint4*)(a + m * 16 + 8)};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint16_t* e = (uint16_t*)&u[h];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = 8 * h + j;
                const float av = bf2f(e[j]);
                acc[c] += g * av;
                e[j] = av > 0.f ? f2bf(g * hw[c]) : (uint16_t)0;
            }
            *(uint4*)(d9 + m * 16 + 8 * h) = u[h];
        }
        acc[16] += g;
    }
    for (int c = 0; c < 17; ++c) s[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 17; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 17) atomicAdd(&ghw[threadIdx.x], s[threadIdx.x][0]);
}

// Adam on the FP32 parameter vector
__global__ void adam_kernel(float* p, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2,
                            float c1, float c2, float eps)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float a = b1 * m1[q] + (1.f - b1) * gq;
        const float b = b2 * m2[q] + (1.f - b2) * gq * gq;
        m1[q] = a;
        m2[q] = b;
        p[q] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
    }
}

// pack FP32 W[27][cin][cout] into MFMA B fragments: [kb][nt][lane][8],
// element (kb, nt, lane, j) = W at k = 32 kb + 8 (lane/16) + j, co = 16 nt +
// lane%16.  flip: the data-gradient weights W'[26 - tap][co][ci] (cin and
// cout swap roles).
__device__ __forceinline__ void pack_w(const float* w, int cin, int cout, int flip, int kb_n, uint16_t* wp)
{
    const int ci_eff = flip ? cout : cin, co_eff = flip ? cin : cout;
    const int nt = co_eff / 16;
    const long n = (long)kb_n * nt * 64 * 8;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int j = (int)(q % 8);
        long r = q / 8;
        const int lane = (int)(r % 64); r /= 64;
        const int t = (int)(r % nt);
        const int kb = (int)(r / nt);
        const int k = kb * 32 + (lane >> 4) * 8 + j;
        const int tap = k / ci_eff, ci = k % ci_eff, co = t * 16 + (lane & 15);
        float v = 0.f;
        if (tap < 27) v = flip ? w[((long)(26 - tap) * cin + co) * cout + ci] : w[((long)tap * cin + ci) * cout + co];
        wp[q] = f2bf(v);
    }
}

// every layer's forward and data-gradient packing in one launch: grid row
// blockIdx.y packs job y (rows past a small job's extent exit at once)
struct PackJob {
    const float* w;
    uint16_t* wp;
    int cin, cout, flip, kb_n;
};
constexpr int kPackJobs = 20;           // two per layer (the network's ten layers)
struct PackJobs {
    PackJob j[kPackJobs];
};
__global__ void pack_all_kernel(PackJobs jobs)
{
    const PackJob& J = jobs.j[blockIdx.y];
    pack_w(J.w, J.cin, J.cout, J.flip, J.kb_n, J.wp);
}

__global__ void fill_f32k(float* p, long n, float v)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) p[q] = v;
}
__global__ void fill_f64k(double* p, long n, double v)
{
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) p[q] = v;
}

// prediction log(1 + phi) -> path length per voxel (zero where mu_t = 0)
__global__ void write_tally_kernel(const float* out, const float* mut, long nvox, double nv, double* tally)
{
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < nvox; m += (long)gridDim.x * blockDim.x)
        tally[m] = mut[m] > 0.f ? fmax(expm1((double)out[m]), 0.0) * nv : 0.0;
}

unsigned grid_of(long n, int per = 256)
{
    long b = (n + per - 1) / per;
    if (b > 16384) b = 16384;
    return (unsigned)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------- network --

struct Layer {
    int cin, cout, level;       // level 0: full grid, 1: /2, 2: /4
    long w_off, b_off;          // offsets in the parameter vector
    int kb_f, kb_b;             // k blocks of the forward / data-gradient GEMM
    uint16_t *wf, *wb;          // packed BF16 weights
};

constexpr int NL = 10;
static_assert(2 * NL == kPackJobs, "one forward and one data-gradient packing per layer");
constexpr int CIN[NL] = {16, 16, 16, 32, 32, 64, 96, 32, 48, 16};
constexpr int COUT[NL] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
constexpr int LEVEL[NL] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0};

}  // namespace

struct tam_dn {
    int dev;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::string err;
    int D[3], H[3], W[3];
    long nv[3];
    Layer L[NL];
    long nparam, head_off;
    float *p, *g, *m1, *m2;           // parameters, gradients, Adam moments
    long step;
    // activations (BF16) and their gradients
    uint16_t *x0, *a[NL], *p1, *p3;
    uint16_t *d[NL], *gc6, *gc8, *gp1, *gp3;
    float *out, *feat;                // FP32 head output; FP32 feature staging (5 x nvox)
    float* wpart;                     // per-block weight-gradient partials of the brick kernel
    long wpart_n;                     // its size in floats
    double* red;                      // [0] loss, [1] input loss
    // dataset
    int capacity;
    uint16_t* ds_feat;                // [slot][nvox][4] BF16 features
    float* ds_tgt;                    // [slot][nvox] FP32 target log(1 + phi_hi)
    float last_ms;
};

namespace {

int dfail(tam_dn* n, int code, const std::string& m)
{
    if (n) n->err = m;
    return code;
}
#define DNCHK(n, call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return dfail((n), TAM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

thread_local std::string g_dn_err;

hipError_t conv_launch(tam_dn* n, const uint16_t* x, int lvl, int cin, const uint16_t* wp, int kb, const float* bias,
                       int cout, int relu, const uint16_t* mask, uint16_t* y, const uint16_t* x2 = nullptr,
                       int c1 = 0)
{
    ConvArgs a{x, wp, bias, mask, y, n->D[lvl], n->H[lvl], n->W[lvl], cin, cout, kb, relu, x2, c1};
    const long nv = n->nv[lvl];
    // 16 x MT-voxel wave tiles: 64 voxels, or 128 for the 48-channel input
    // layer (its long k loop amortises the extra registers); every level's
    // voxel count is a multiple of 512
    if (nv % 512 != 0 || kb != (27 * cin + 31) / 32) return hipErrorInvalidValue;
    const unsigned b8 = (unsigned)(nv / (4 * 16 * 8)), b4 = (unsigned)(nv / (4 * 16 * 4));
#define CV(CI, NT_, MT_) \
    hipLaunchKernelGGL((conv_kernel<CI, NT_, MT_>), dim3(MT_ == 8 ? b8 : b4), dim3(256), 0, n->stream, a)
    const int D = n->D[lvl], H = n->H[lvl], W = n->W[lvl];
    if (W % 16 == 0 && H % 8 == 0 && D % 4 == 0) {
        // LDS-staged input for the layers with 16 or 32 output channels (48
        // outputs as NT = 3 needs ~260 VGPRs there: the global form is faster)
        const dim3 grid((unsigned)((W / 16) * (H / 8) * (D / 4)));
#define CL(CI, NT_) hipLaunchKernelGGL((conv_lds_kernel<CI, NT_>), grid, dim3(256), 0, n->stream, a)
        switch (cin * 1000 + cout) {
        case 16016: CL(16, 1); return hipGetLastError();
        case 16032: CL(16, 2); return hipGetLastError();
        case 32032: CL(32, 2); return hipGetLastError();
        case 96032: CL(96, 2); return hipGetLastError();
        case 48016: CL(48, 1); return hipGetLastError();
        case 32016: CL(32, 1); return hipGetLastError();
        }
#undef CL
    }
    if (x2) return hipErrorInvalidValue;  // a concatenated input is read by the LDS form only
    switch (cin * 1000 + cout) {
    case 16016: CV(16, 1, 4); break;      // layers 0, 1, 9 and the data gradients of 0, 1, 9
    case 16032: CV(16, 2, 4); break;      // layer 2
    case 32032: CV(32, 2, 4); break;      // layers 3, 7 and their data gradients
    case 32064: CV(32, 4, 4); break;      // layer 4
    case 64064: CV(64, 4, 4); break;      // layer 5 and its data gradient
    case 96032: CV(96, 2, 4); break;      // layer 6
    case 48016: CV(48, 1, 8); break;      // layer 8
    case 32016: CV(32, 1, 4); break;      // data gradient of layer 2
    case 64032: CV(64, 2, 4); break;      // data gradient of layer 4
    case 32096: CV(32, 6, 4); break;      // data gradient of layer 6
    case 16048: CV(16, 3, 4); break;      // data gradient of layer 8
    default: return hipErrorInvalidValue;
    }
#undef CV
    return hipGetLastError();
}

// The brick weight-gradient launch of a layer: grid rows (input-channel
// tiles x dz groups), dz groups, blocks per row and partial floats per block;
// false where the layer keeps the segment kernels (grid edges not multiples of
// 16 / 8 / 4).
struct BrickPlan {
    int ny, ndz;
    long nb, per;
};
bool brick_plan(int D, int H, int W, int cin, int cout, BrickPlan* bp)
{
    if (W % 16 != 0 || H % 8 != 0 || D % 4 != 0 || cin % 16 != 0) return false;
    if (cout != 16 && cout != 32 && cout != 64) return false;
    bp->ndz = (cout == 16 && cin <= 48) ? 1 : 3;           // all 27 taps' sums in registers when they fit
    bp->ny = (cin / 16) * bp->ndz;
    const long bricks = (long)(W / 16) * (H / 8) * (D / 4);
    long nb = 1024 / bp->ny;
    if (nb < 1) nb = 1;
    bp->nb = nb < bricks ? nb : bricks;
    bp->per = (27 / bp->ndz) * 16L * cout + cout;
    return true;
}

// weight gradient of one layer into gw; the brick form also sums the bias
// gradient into gb and says so in *bias_done
hipError_t wgrad_launch(tam_dn* n, const uint16_t* x, const uint16_t* dy, int lvl, int cin, int cout, float* gw,
                        float* gb, bool* bias_done, const uint16_t* x2 = nullptr, int c1 = 0)
{
    *bias_done = false;
    BrickPlan bp;
    if (brick_plan(n->D[lvl], n->H[lvl], n->W[lvl], cin, cout, &bp) && n->wpart &&
        (long)bp.ny * bp.nb * bp.per <= n->wpart_n) {
        WgradbArgs b{x, dy, n->wpart, n->D[lvl], n->H[lvl], n->W[lvl], cin, cout,
                     (long)(n->W[lvl] / 16) * (n->H[lvl] / 8) * (n->D[lvl] / 4), x2, c1};
        const dim3 grid((unsigned)bp.nb, (unsigned)bp.ny);
        if (cout == 16 && bp.ndz == 1) hipLaunchKernelGGL((wgradb_kernel<1, 1>), grid, dim3(256), 0, n->stream, b);
        else if (cout == 16) hipLaunchKernelGGL((wgradb_kernel<1, 3>), grid, dim3(256), 0, n->stream, b);
        else if (cout == 32) hipLaunchKernelGGL((wgradb_kernel<2, 3>), grid, dim3(256), 0, n->stream, b);
        else hipLaunchKernelGGL((wgradb_kernel<4, 3>), grid, dim3(256), 0, n->stream, b);
        // ~1024 reduction blocks: each weight's nb partials split over ksplit of them
        const unsigned qb = grid_of(bp.ny * bp.per);
        long ks = 1024 / (long)qb;
        if (ks < 1) ks = 1;
        if (ks > bp.nb) ks = bp.nb;
        hipLaunchKernelGGL(wgradb_reduce_kernel, dim3(qb, (unsigned)ks), dim3(256), 0, n->stream, n->wpart,
                           (int)bp.nb, bp.ny, bp.ndz, cin, cout, gw, gb);
        *bias_done = true;
        return hipGetLastError();
    }
    if (x2) return hipErrorInvalidValue;  // a concatenated input is read by the brick form only
    WgradArgs a{x, dy, gw, n->D[lvl], n->H[lvl], n->W[lvl], cin, cout, 0, 0};
    a.blocks = n->nv[lvl] / 32;
    // rows of whole 32-voxel segments: one workgroup per (dz, dy) covers the three dx taps, ~288 per (dz, dy);
    // otherwise one per tap, ~96 per tap (~2600 workgroups in all either way)
    // (the 64 x 64 layer keeps one workgroup per tap: three taps' accumulators leave it one wave per SIMD)
    const bool rows = n->W[lvl] % 32 == 0 && (cin / 16) * (cout / 16) <= 12;
    const long wgs = rows ? 288 : 96;
    long per = (a.blocks + wgs - 1) / wgs;
    per = (per + 3) / 4 * 4;
    if (per < 4) per = 4;
    a.per_wg = per;
    const dim3 grid((unsigned)((a.blocks + per - 1) / per), rows ? 9 : 27);
#define WG(CI, CO)                                                                                     \
    do {                                                                                               \
        if (rows) hipLaunchKernelGGL((wgrad3_kernel<CI, CO>), grid, dim3(256), 0, n->stream, a);       \
        else hipLaunchKernelGGL((wgrad_kernel<CI, CO>), grid, dim3(256), 0, n->stream, a);             \
    } while (0)
    const int key = (cin / 16) * 16 + cout / 16;
    switch (key) {
    case 1 * 16 + 1: WG(1, 1); break;
    case 1 * 16 + 2: WG(1, 2); break;
    case 2 * 16 + 2: WG(2, 2); break;
    case 2 * 16 + 4: WG(2, 4); break;
    case 4 * 16 + 4: WG(4, 4); break;
    case 6 * 16 + 2: WG(6, 2); break;
    case 3 * 16 + 1: WG(3, 1); break;
    default: return hipErrorInvalidValue;
    }
#undef WG
    return hipGetLastError();
}

hipError_t repack(tam_dn* n)
{
    PackJobs jobs;
    unsigned gx = 1;
    for (int l = 0; l < NL; ++l) {
        const Layer& L = n->L[l];
        const float* w = n->p + L.w_off;
        jobs.j[2 * l] = PackJob{w, L.wf, L.cin, L.cout, 0, L.kb_f};
        jobs.j[2 * l + 1] = PackJob{w, L.wb, L.cin, L.cout, 1, L.kb_b};
        gx = std::max(gx, grid_of((long)L.kb_f * L.cout * 32));
        gx = std::max(gx, grid_of((long)L.kb_b * L.cin * 32));
    }
    hipLaunchKernelGGL(pack_all_kernel, dim3(gx, kPackJobs), dim3(256), 0, n->stream, jobs);
    return hipGetLastError();
}

// forward from n->x0; with tgt: the L1 loss summed into red[0] (input's in red[1])
hipError_t forward(tam_dn* n, const float* tgt)
{
    hipError_t e;
    auto conv = [&](int l, const uint16_t* in, uint16_t* o) {
        const Layer& L = n->L[l];
        return conv_launch(n, in, L.level, L.cin, L.wf, L.kb_f, n->p + L.b_off, L.cout, 1, nullptr, o);
    };
    if ((e = conv(0, n->x0, n->a[0])) != hipSuccess) return e;
    if ((e = conv(1, n->a[0], n->a[1])) != hipSuccess) return e;
    hipLaunchKernelGGL(pool_kernel, dim3(grid_of(n->nv[1] * 16 / 8)), dim3(256), 0, n->stream, n->a[1], n->D[0], n->H[0],
                       n->W[0], 16, n->p1);
    if ((e = conv(2, n->p1, n->a[2])) != hipSuccess) return e;
    if ((e = conv(3, n->a[2], n->a[3])) != hipSuccess) return e;
    hipLaunchKernelGGL(pool_kernel, dim3(grid_of(n->nv[2] * 32 / 8)), dim3(256), 0, n->stream, n->a[3], n->D[1], n->H[1],
                       n->W[1], 32, n->p3);
    if ((e = conv(4, n->p3, n->a[4])) != hipSuccess) return e;
    if ((e = conv(5, n->a[4], n->a[5])) != hipSuccess) return e;
    // layers 6 and 8 read cat(upsample2(coarse), skip) in their input staging
    auto conv_cat = [&](int l, const uint16_t* coarse, int c1, const uint16_t* skip, uint16_t* o) {
        const Layer& L = n->L[l];
        return conv_launch(n, coarse, L.level, L.cin, L.wf, L.kb_f, n->p + L.b_off, L.cout, 1, nullptr, o, skip, c1);
    };
    if ((e = conv_cat(6, n->a[5], 64, n->a[3], n->a[6])) != hipSuccess) return e;
    if ((e = conv(7, n->a[6], n->a[7])) != hipSuccess) return e;
    if ((e = conv_cat(8, n->a[7], 32, n->a[1], n->a[8])) != hipSuccess) return e;
    if ((e = conv(9, n->a[8], n->a[9])) != hipSuccess) return e;
    if (tgt) hipLaunchKernelGGL(fill_f64k, dim3(1), dim3(64), 0, n->stream, n->red, 2L, 0.0);
    // with a target every workgroup ends in atomics onto two addresses: at most
    // 1024 workgroups then; without one, a workgroup per 256 voxels (up to
    // grid_of's cap) so the loads of many voxels are in flight at once
    const unsigned hg = grid_of(n->nv[0]);
    const unsigned hb = tgt && hg > 1024u ? 1024u : hg;
    hipLaunchKernelGGL(head_kernel, dim3(hb), dim3(256), 0, n->stream, n->x0, n->a[9],
                       n->p + n->head_off, n->nv[0], n->out, tgt, n->red, n->red + 1);
    return hipGetLastError();
}

// backward after forward(tgt): gradients of the mean L1 into n->g
hipError_t backward(tam_dn* n, const float* tgt)
{
    hipError_t e;
    hipLaunchKernelGGL(fill_f32k, dim3(grid_of(n->nparam)), dim3(256), 0, n->stream, n->g, n->nparam, 0.f);
    const unsigned hb = grid_of(n->nv[0]) < 1024u ? grid_of(n->nv[0]) : 1024u;
    hipLaunchKernelGGL(head_bwd_kernel, dim3(hb), dim3(256), 0, n->stream, n->out, tgt, n->a[9],
                       n->p + n->head_off, n->nv[0], n->d[9], n->g + n->head_off);
    auto grads = [&](int l, const uint16_t* in, const uint16_t* dl, const uint16_t* x2 = nullptr, int c1 = 0) {
        const Layer& L = n->L[l];
        bool bias_done = false;
        hipError_t r = wgrad_launch(n, in, dl, L.level, L.cin, L.cout, n->g + L.w_off, n->g + L.b_off, &bias_done,
                                    x2, c1);
        if (r != hipSuccess || bias_done) return r;
        // at most 1024 blocks: every block ends in `cout` atomics onto the same few addresses, which serialise
        if (L.cout % 8 != 0 || L.cout > 2048) return hipErrorInvalidValue;
        unsigned nb = grid_of(n->nv[L.level], 256 / (L.cout / 8));
        hipLaunchKernelGGL(bgrad_kernel, dim3(nb < 1024u ? nb : 1024u), dim3(256), 0, n->stream, dl,
                           n->nv[L.level], L.cout, n->g + L.b_off);
        return hipGetLastError();
    };
    // data gradient of layer l into `o`, masked by `mask` (the ReLU output feeding l), if any
    auto dgrad = [&](int l, const uint16_t* dl, const uint16_t* mask, uint16_t* o) {
        const Layer& L = n->L[l];
        return conv_launch(n, dl, L.level, L.cout, L.wb, L.kb_b, nullptr, L.cin, 0, mask, o);
    };
    if ((e = grads(9, n->a[8], n->d[9])) != hipSuccess) return e;
    if ((e = dgrad(9, n->d[9], n->a[8], n->d[8])) != hipSuccess) return e;
    if ((e = grads(8, n->a[7], n->d[8], n->a[1], 32)) != hipSuccess) return e;
    if ((e = dgrad(8, n->d[8], nullptr, n->gc8)) != hipSuccess) return e;
    hipLaunchKernelGGL(upcat_bwd_kernel, dim3(grid_of(n->nv[1] * 32 / 8)), dim3(256), 0, n->stream, n->gc8, 48, 32,
                       n->a[7], n->D[0], n->H[0], n->W[0], n->d[7]);
    if ((e = grads(7, n->a[6], n->d[7])) != hipSuccess) return e;
    if ((e = dgrad(7, n->d[7], n->a[6], n->d[6])) != hipSuccess) return e;
    if ((e = grads(6, n->a[5], n->d[6], n->a[3], 64)) != hipSuccess) return e;
    if ((e = dgrad(6, n->d[6], nullptr, n->gc6)) != hipSuccess) return e;
    hipLaunchKernelGGL(upcat_bwd_kernel, dim3(grid_of(n->nv[2] * 64 / 8)), dim3(256), 0, n->stream, n->gc6, 96, 64,
                       n->a[5], n->D[1], n->H[1], n->W[1], n->d[5]);
    if ((e = grads(5, n->a[4], n->d[5])) != hipSuccess) return e;
    if ((e = dgrad(5, n->d[5], n->a[4], n->d[4])) != hipSuccess) return e;
    if ((e = grads(4, n->p3, n->d[4])) != hipSuccess) return e;
    if ((e = dgrad(4, n->d[4], nullptr, n->gp3)) != hipSuccess) return e;
    hipLaunchKernelGGL(skip_pool_bwd_kernel, dim3(grid_of(n->nv[1] * 32 / 8)), dim3(256), 0, n->stream, n->gc6, 96, 64,
                       n->gp3, n->a[3], n->D[1], n->H[1], n->W[1], 32, n->d[3]);
    if ((e = grads(3, n->a[2], n->d[3])) != hipSuccess) return e;
    if ((e = dgrad(3, n->d[3], n->a[2], n->d[2])) != hipSuccess) return e;
    if ((e = grads(2, n->p1, n->d[2])) != hipSuccess) return e;
    if ((e = dgrad(2, n->d[2], nullptr, n->gp1)) != hipSuccess) return e;
    hipLaunchKernelGGL(skip_pool_bwd_kernel, dim3(grid_of(n->nv[0] * 16 / 8)), dim3(256), 0, n->stream, n->gc8, 48, 32,
                       n->gp1, n->a[1], n->D[0], n->H[0], n->W[0], 16, n->d[1]);
    if ((e = grads(1, n->a[0], n->d[1])) != hipSuccess) return e;
    if ((e = dgrad(1, n->d[1], n->a[0], n->d[0])) != hipSuccess) return e;
    if ((e = grads(0, n->x0, n->d[0])) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t adam(tam_dn* n, float lr)
{
    ++n->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)n->step));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)n->step));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_of(n->nparam)), dim3(256), 0, n->stream, n->p, n->g, n->m1, n->m2,
                       n->nparam, lr, b1, b2, c1, c2, 1e-8f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return repack(n);
}

struct DevScope {
    int prev = -1;
    explicit DevScope(int d)
    {
        hipGetDevice(&prev);
        if (prev != d) hipSetDevice(d);
    }
    ~DevScope()
    {
        int cur = -1;
        hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) hipSetDevice(prev);
    }
};

float ms_between(tam_dn* n)
{
    float ms = 0.f;
    hipEventElapsedTime(&ms, n->ev0, n->ev1);
    return ms;
}

}  // namespace

extern "C" {

int tam_dn_create(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t capacity, uint64_t seed, tam_dn** out)
{
    if (!out) return TAM_ERR_ARG;
    *out = nullptr;
    if (nx < 32 || ny < 32 || nz < 32 || nx % 32 || ny % 32 || nz % 32 || capacity < 0) {
        g_dn_err = "grid edges must be positive multiples of 32";
        return TAM_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_dn_err = "no HIP device";
        return TAM_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
        g_dn_err = "device ordinal out of range";
        return TAM_ERR_ARG;
    }
    tam_dn* n = new (std::nothrow) tam_dn();
    if (!n) return TAM_ERR_NOMEM;
    n->dev = device;
    DevScope ds(device);
    for (int l = 0; l < 3; ++l) {
        n->W[l] = nx >> l;
        n->H[l] = ny >> l;
        n->D[l] = nz >> l;
        n->nv[l] = (long)n->W[l] * n->H[l] * n->D[l];
    }
    long off = 0;
    for (int l = 0; l < NL; ++l) {
        Layer& L = n->L[l];
        L.cin = CIN[l];
        L.cout = COUT[l];
        L.level = LEVEL[l];
        L.w_off = off;
        off += 27L * L.cin * L.cout;
        L.b_off = off;
        off += L.cout;
        L.kb_f = (27 * L.cin + 31) / 32;
        L.kb_b = (27 * L.cout + 31) / 32;
    }
    n->head_off = off;
    off += 17;
    n->nparam = off;
    n->capacity = capacity;
    hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&n->ev0);
    if (e == hipSuccess) e = hipEventCreate(&n->ev1);
    auto al = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes > 0 ? bytes : 16);
    };
    al((void**)&n->p, sizeof(float) * n->nparam);
    al((void**)&n->g, sizeof(float) * n->nparam);
    al((void**)&n->m1, sizeof(float) * n->nparam);
    al((void**)&n->m2, sizeof(float) * n->nparam);
    for (int l = 0; l < NL; ++l) {
        Layer& L = n->L[l];
        al((void**)&L.wf, 2ull * L.kb_f * 32 * L.cout);
        al((void**)&L.wb, 2ull * L.kb_b * 32 * L.cin);
        al((void**)&n->a[l], 2ull * n->nv[L.level] * L.cout);
        al((void**)&n->d[l], 2ull * n->nv[L.level] * L.cout);
    }
    al((void**)&n->x0, 2ull * n->nv[0] * 16);
    al((void**)&n->p1, 2ull * n->nv[1] * 16);
    al((void**)&n->p3, 2ull * n->nv[2] * 32);
    al((void**)&n->gc6, 2ull * n->nv[1] * 96);
    al((void**)&n->gc8, 2ull * n->nv[0] * 48);
    al((void**)&n->gp1, 2ull * n->nv[1] * 16);
    al((void**)&n->gp3, 2ull * n->nv[2] * 32);
    al((void**)&n->out, 4ull * n->nv[0]);
    al((void**)&n->feat, 4ull * 5 * n->nv[0]);
    al((void**)&n->red, 16);
    n->wpart_n = 0;
    for (int l = 0; l < NL; ++l) {
        const Layer& L = n->L[l];
        BrickPlan bp;
        if (brick_plan(n->D[L.level], n->H[L.level], n->W[L.level], L.cin, L.cout, &bp))
            n->wpart_n = std::max(n->wpart_n, (long)bp.ny * bp.nb * bp.per);
    }
    if (n->wpart_n > 0) al((void**)&n->wpart, sizeof(float) * n->wpart_n);
    al((void**)&n->ds_feat, 8ull * n->nv[0] * (capacity > 0 ? capacity : 1));
    al((void**)&n->ds_tgt, 4ull * n->nv[0] * (capacity > 0 ? capacity : 1));
    if (e != hipSuccess) {
        g_dn_err = std::string("denoiser allocation: ") + hipGetErrorString(e);
        tam_dn_destroy(n);
        return e == hipErrorOutOfMemory ? TAM_ERR_NOMEM : TAM_ERR_HIP;
    }
    // He initialisation (input layer: 4 real channels; the 12 padding
    // channels always see zeros), zero biases, zero head (output = input)
    std::vector<float> h((size_t)n->nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NL; ++l) {
        const Layer& L = n->L[l];
        const int fan = 27 * (l == 0 ? 4 : L.cin);
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan));
        for (int t = 0; t < 27; ++t)
            for (int ci = 0; ci < L.cin; ++ci)
                for (int co = 0; co < L.cout; ++co)
                    h[(size_t)(L.w_off + ((long)t * L.cin + ci) * L.cout + co)] = (l == 0 && ci >= 4) ? 0.f : nd(rng);
    }
    e = hipMemcpy(n->p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->m1, 0, sizeof(float) * n->nparam);
    if (e == hipSuccess) e = hipMemset(n->m2, 0, sizeof(float) * n->nparam);
    if (e == hipSuccess) e = hipMemset(n->x0, 0, 2ull * n->nv[0] * 16);
    if (e == hipSuccess) e = repack(n);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) {
        g_dn_err = std::string("denoiser init: ") + hipGetErrorString(e);
        tam_dn_destroy(n);
        return TAM_ERR_HIP;
    }
    *out = n;
    return TAM_OK;
}

void tam_dn_destroy(tam_dn* n)
{
    if (!n) return;
    DevScope ds(n->dev);
    if (n->stream) hipStreamSynchronize(n->stream);
    std::vector<void*> ptrs = {n->p, n->g, n->m1, n->m2, n->x0, n->p1, n->p3, n->gc6, n->gc8,
                               n->gp1, n->gp3, n->out, n->feat, n->red, n->ds_feat, n->ds_tgt, n->wpart};
    for (int l = 0; l < NL; ++l) {
        ptrs.push_back(n->L[l].wf);
        ptrs.push_back(n->L[l].wb);
        ptrs.push_back(n->a[l]);
        ptrs.push_back(n->d[l]);
    }
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (n->ev0) hipEventDestroy(n->ev0);
    if (n->ev1) hipEventDestroy(n->ev1);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
}

const char* tam_dn_last_error(const tam_dn* n) { return n ? n->err.c_str() : g_dn_err.c_str(); }

int64_t tam_dn_num_params(const tam_dn* n) { return n ? n->nparam : -1; }

int tam_dn_get_params(tam_dn* n, float* out)
{
    if (!n || !out) return TAM_ERR_ARG;
    DevScope ds(n->dev);
    DNCHK(n, hipMemcpyAsync(out, n->p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    DNCHK(n, hipStreamSynchronize(n->stream));
    return TAM_OK;
}

int tam_dn_set_params(tam_dn* n, const float* in)
{
    if (!n || !in) return TAM_ERR_ARG;
    DevScope ds(n->dev);
    DNCHK(n, hipMemcpyAsync(n->p, in, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->stream));
    DNCHK(n, hipMemsetAsync(n->m1, 0, sizeof(float) * n->nparam, n->stream));
    DNCHK(n, hipMemsetAsync(n->m2, 0, sizeof(float) * n->nparam, n->stream));
    n->step = 0;
    DNCHK(n, repack(n));
    DNCHK(n, hipStreamSynchronize(n->stream));
    return TAM_OK;
}

int tam_dn_store(tam_dn* n, int32_t slot, const tam_ctx* lo, double npk_lo, const tam_ctx* hi, double npk_hi)
{
    if (!n || !lo || !hi) return TAM_ERR_ARG;
    if (slot < 0 || slot >= n->capacity) return dfail(n, TAM_ERR_ARG, "slot out of range");
    if (lo->p.nx != n->W[0] || lo->p.ny != n->H[0] || lo->p.nz != n->D[0] || hi->p.nx != n->W[0] ||
        hi->p.ny != n->H[0] || hi->p.nz != n->D[0] || lo->dev != n->dev || hi->dev != n->dev)
        return dfail(n, TAM_ERR_ARG, "contexts must match the network's grid and device");
    if (!(npk_lo > 0 && npk_hi > 0)) return dfail(n, TAM_ERR_ARG, "packet counts must be positive");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    const double vcm = lo->p.volume_voxel * 1e6;
    // the contexts' streams finish their transport before this stream reads
    DNCHK(n, hipStreamSynchronize(lo->stream));
    DNCHK(n, hipStreamSynchronize(hi->stream));
    hipLaunchKernelGGL(ctx_features_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, lo->tally, lo->mut, lo->mat,
                       lo->mtab, nv, 1.0 / (npk_lo * vcm), n->feat);
    hipLaunchKernelGGL(store_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->feat, hi->tally,
                       1.0 / (npk_hi * vcm), nv, n->ds_feat + (long)slot * nv * 4, n->ds_tgt + (long)slot * nv);
    DNCHK(n, hipGetLastError());
    DNCHK(n, hipStreamSynchronize(n->stream));
    return TAM_OK;
}

int tam_dn_store_host(tam_dn* n, int32_t slot, const float* features, const float* target)
{
    if (!n || !features || !target) return TAM_ERR_ARG;
    if (slot < 0 || slot >= n->capacity) return dfail(n, TAM_ERR_ARG, "slot out of range");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    DNCHK(n, hipMemcpyAsync(n->feat, features, sizeof(float) * 4 * nv, hipMemcpyHostToDevice, n->stream));
    DNCHK(n, hipMemcpyAsync(n->feat + 4 * nv, target, sizeof(float) * nv, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(store_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->feat, (const double*)nullptr,
                       0.0, nv, n->ds_feat + (long)slot * nv * 4, n->ds_tgt + (long)slot * nv);
    DNCHK(n, hipGetLastError());
    DNCHK(n, hipStreamSynchronize(n->stream));
    return TAM_OK;
}

int tam_dn_train(tam_dn* n, int32_t nslots, int32_t steps, double lr, uint64_t seed, float* losses)
{
    if (!n) return TAM_ERR_ARG;
    if (nslots < 1 || nslots > n->capacity || steps < 0 || !(lr > 0))
        return dfail(n, TAM_ERR_ARG, "need 1 <= nslots <= capacity, steps >= 0, lr > 0");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    std::mt19937_64 rng(seed);
    std::vector<double> red((size_t)steps > 0 ? (size_t)steps * 2 : 2);
    double* dred = nullptr;
    DNCHK(n, hipMalloc((void**)&dred, sizeof(double) * 2 * (size_t)(steps > 0 ? steps : 1)));
    DNCHK(n, hipEventRecord(n->ev0, n->stream));
    hipError_t e = hipSuccess;
    for (int s = 0; s < steps && e == hipSuccess; ++s) {
        const int slot = (int)(rng() % (uint64_t)nslots);
        const float* tgt = n->ds_tgt + (long)slot * nv;
        hipLaunchKernelGGL(pack_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream,
                           n->ds_feat + (long)slot * nv * 4, nv, n->x0);
        e = forward(n, tgt);
        if (e == hipSuccess) e = hipMemcpyAsync(dred + 2 * s, n->red, 16, hipMemcpyDeviceToDevice, n->stream);
        if (e == hipSuccess) e = backward(n, tgt);
        if (e == hipSuccess) e = adam(n, (float)lr);
    }
    if (e == hipSuccess) e = hipEventRecord(n->ev1, n->stream);
    if (e == hipSuccess && steps > 0)
        e = hipMemcpyAsync(red.data(), dred, sizeof(double) * 2 * steps, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    hipFree(dred);
    if (e != hipSuccess) return dfail(n, TAM_ERR_HIP, std::string("training: ") + hipGetErrorString(e));
    n->last_ms = ms_between(n);
    if (losses)
        for (int s = 0; s < steps; ++s) losses[s] = (float)(red[2 * (size_t)s] / (double)nv);
    return TAM_OK;
}

int tam_dn_eval(tam_dn* n, int32_t slot, double* l1_pred, double* l1_input)
{
    if (!n) return TAM_ERR_ARG;
    if (slot < 0 || slot >= n->capacity) return dfail(n, TAM_ERR_ARG, "slot out of range");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    const float* tgt = n->ds_tgt + (long)slot * nv;
    hipLaunchKernelGGL(pack_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->ds_feat + (long)slot * nv * 4,
                       nv, n->x0);
    DNCHK(n, hipEventRecord(n->ev0, n->stream));
    DNCHK(n, forward(n, tgt));
    DNCHK(n, hipEventRecord(n->ev1, n->stream));
    double r[2];
    DNCHK(n, hipMemcpyAsync(r, n->red, 16, hipMemcpyDeviceToHost, n->stream));
    DNCHK(n, hipStreamSynchronize(n->stream));
    n->last_ms = ms_between(n);
    if (l1_pred) *l1_pred = r[0] / (double)nv;
    if (l1_input) *l1_input = r[1] / (double)nv;
    return TAM_OK;
}

int tam_dn_predict_slot(tam_dn* n, int32_t slot, float* pred, float* target, float* input)
{
    if (!n) return TAM_ERR_ARG;
    if (slot < 0 || slot >= n->capacity) return dfail(n, TAM_ERR_ARG, "slot out of range");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    hipLaunchKernelGGL(pack_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->ds_feat + (long)slot * nv * 4,
                       nv, n->x0);
    DNCHK(n, forward(n, nullptr));
    if (pred) DNCHK(n, hipMemcpyAsync(pred, n->out, sizeof(float) * nv, hipMemcpyDeviceToHost, n->stream));
    if (target)
        DNCHK(n, hipMemcpyAsync(target, n->ds_tgt + (long)slot * nv, sizeof(float) * nv, hipMemcpyDeviceToHost,
                                n->stream));
    if (input) {
        std::vector<uint16_t> f((size_t)nv * 4);
        DNCHK(n, hipMemcpyAsync(f.data(), n->ds_feat + (long)slot * nv * 4, 8 * (size_t)nv, hipMemcpyDeviceToHost,
                                n->stream));
        DNCHK(n, hipStreamSynchronize(n->stream));
        for (long m = 0; m < nv; ++m) {
            const uint32_t u = (uint32_t)f[(size_t)m * 4] << 16;
            float v;
            std::memcpy(&v, &u, 4);
            input[m] = v;
        }
    }
    DNCHK(n, hipStreamSynchronize(n->stream));
    return TAM_OK;
}

int tam_dn_forward_host(tam_dn* n, const float* features, float* out)
{
    if (!n || !features || !out) return TAM_ERR_ARG;
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    DNCHK(n, hipMemcpyAsync(n->feat, features, sizeof(float) * 4 * nv, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(pack_input_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->feat, nv, n->x0);
    DNCHK(n, hipEventRecord(n->ev0, n->stream));
    DNCHK(n, forward(n, nullptr));
    DNCHK(n, hipEventRecord(n->ev1, n->stream));
    DNCHK(n, hipMemcpyAsync(out, n->out, sizeof(float) * nv, hipMemcpyDeviceToHost, n->stream));
    DNCHK(n, hipStreamSynchronize(n->stream));
    n->last_ms = ms_between(n);
    return TAM_OK;
}

int tam_dn_gradients_host(tam_dn* n, const float* features, const float* target, float* grads, double* loss)
{
    if (!n || !features || !target || !grads) return TAM_ERR_ARG;
    if (n->capacity < 1) return dfail(n, TAM_ERR_ARG, "needs one dataset slot");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    const int rc = tam_dn_store_host(n, 0, features, target);
    if (rc != TAM_OK) return rc;
    const float* tgt = n->ds_tgt;
    hipLaunchKernelGGL(pack_slot_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->ds_feat, nv, n->x0);
    DNCHK(n, forward(n, tgt));
    DNCHK(n, backward(n, tgt));
    double r[2];
    DNCHK(n, hipMemcpyAsync(r, n->red, 16, hipMemcpyDeviceToHost, n->stream));
    DNCHK(n, hipMemcpyAsync(grads, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    DNCHK(n, hipStreamSynchronize(n->stream));
    if (loss) *loss = r[0] / (double)nv;
    return TAM_OK;
}

int tam_dn_apply(tam_dn* n, tam_ctx* c, double npk)
{
    if (!n || !c) return TAM_ERR_ARG;
    if (c->p.nx != n->W[0] || c->p.ny != n->H[0] || c->p.nz != n->D[0] || c->dev != n->dev)
        return dfail(n, TAM_ERR_ARG, "the context must match the network's grid and device");
    if (!(npk > 0)) return dfail(n, TAM_ERR_ARG, "packet count must be positive");
    DevScope ds(n->dev);
    const long nv = n->nv[0];
    const double vcm = c->p.volume_voxel * 1e6;
    DNCHK(n, hipStreamSynchronize(c->stream));
    DNCHK(n, hipEventRecord(n->ev0, n->stream));
    hipLaunchKernelGGL(ctx_input_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, c->tally, c->mut, c->mat,
                       c->mtab, nv, 1.0 / (npk * vcm), n->x0);
    DNCHK(n, forward(n, nullptr));
    hipLaunchKernelGGL(write_tally_kernel, dim3(grid_of(nv)), dim3(256), 0, n->stream, n->out, c->mut, nv, npk * vcm,
                       c->tally);
    DNCHK(n, hipGetLastError());
    DNCHK(n, hipEventRecord(n->ev1, n->stream));
    DNCHK(n, hipStreamSynchronize(n->stream));
    n->last_ms = ms_between(n);
    return TAM_OK;
}

double tam_dn_last_ms(const tam_dn* n) { return n ? (double)n->last_ms : 0.0; }

}  // extern "C"
