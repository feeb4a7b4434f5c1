// unet.hip - learned denoiser: a 5-level convolutional U-Net on BF16 MFMA.
//
// No counterpart in the reference renderer, whose learned denoiser is a prebuilt library
// called from PathTracingRenderer.cpp:244-268; the network, its kernels and its training
// are this file's own. The layer table and the parameter layout are stated in
// include/gknx_rt.h next to gk_unet_create.
//
// Data layout: activations are BF16 NHWC with channel counts padded to multiples of 16, so
// one pixel's 8-channel group is a 16-byte load. Each convolution is an implicit GEMM
//     out^T[cout][pixel] = W[cout][k] x im2col[k][pixel],   k = tap * cin + ci,
// on v_mfma_f32_16x16x32_bf16 (FP32 accumulation). The bias is column k = 9*cin of W,
// whose im2col row is 1, so it is part of the GEMM. The operand gather fuses the zero
// padding, the nearest-neighbour upsampling of the first input and the skip
// concatenation; no im2col or concatenated tensor is written.
//
// Operand lane layouts of v_mfma_f32_16x16x32_bf16 (64 lanes):
//   A (16 x 32): lane & 15 = row, k = (lane >> 4) * 8 + e for the 8 bf16 of the lane
//   B (32 x 16): lane & 15 = column, k = (lane >> 4) * 8 + e
//   C (16 x 16): lane holds column lane & 15 and rows 4 * (lane >> 4) + e, e = 0..3
// A lane's 8 consecutive k never cross a tap because cin is a multiple of 8.
//
// Kernels
//   k_conv    forward convolution (also the input gradient, with flipped weights)
//   k_wgrad   weight gradient: dW[cout][k] = sum_pixels g[pixel][cout] * im2col[k][pixel]
//   k_pool / k_pool_bwd / k_upsum / k_relu_mask / k_loss / k_adam / k_wflip / k_pack /
//   k_unpack / k_frame_input: elementwise and reductions
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gk_post.hpp"
#include "gk_scene.hpp"

namespace gk {
namespace unet {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr uint32_t kIn = GK_UNET_IN_CHANNELS;  // input features
constexpr uint32_t kInPad = 16;                // padded input channels
constexpr uint32_t kLayers = GK_UNET_LAYERS;

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ inline bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }
__device__ inline bf16x8 one8() { return __builtin_bit_cast(bf16x8, make_uint4(0x3f80u, 0u, 0u, 0u)); }
__device__ inline bf16x8 pack8(const uint16_t (&v)[8]) {
    return __builtin_bit_cast(bf16x8, make_uint4(uint32_t(v[0]) | (uint32_t(v[1]) << 16), uint32_t(v[2]) | (uint32_t(v[3]) << 16),
                                                 uint32_t(v[4]) | (uint32_t(v[5]) << 16), uint32_t(v[6]) | (uint32_t(v[7]) << 16)));
}
__device__ inline void st4(uint16_t* p, float a, float b, float c, float d) {
    *reinterpret_cast<uint2*>(p) = make_uint2(uint32_t(f2bf(a)) | (uint32_t(f2bf(b)) << 16),
                                              uint32_t(f2bf(c)) | (uint32_t(f2bf(d)) << 16));
}

// ---------------------------------------------------------------- convolution
struct ConvK {
    const uint16_t* a;  // first input, NHWC with ca channels; at half resolution when a_up
    const uint16_t* b;  // skip input, NHWC with cb channels (cb = 0: none)
    const uint16_t* w;  // [cout][kpad]
    uint16_t* out;      // channels [0, split), stride split
    uint16_t* out2;     // channels [split, cout), stride cout - split
    float* outf;        // when set: channels 0..3 as one float4 per pixel
    uint32_t ca, cb, cin, a_up;
    uint32_t W, H, tiles_y;
    uint32_t kpad, ktap;  // ktap = 9 * cin
    uint32_t cout, split, relu, bias;
};

// One work-group = 4 waves = a 16 x 16 pixel tile of one image and MT*16 output channels;
// a wave owns 4 rows of 16 pixels (4 B tiles) and MT A tiles. MODE 0 gathers operands from
// memory. MODE >= 1 (staged): the work-group first copies its input tile into LDS (zero
// outside the image) and the 9 taps read their operands from there instead of 9 gathers
// from memory: 18 x 18 pixels, stored as one plane of 16-byte units per 8-channel group
// (a B operand's 8 lanes of one LDS cycle then read 128 contiguous bytes: no bank
// conflicts), and for an upsampled first input its 10 x 10 half-resolution pixels in planes
// of their own (W and H are even, so a pixel lies in the image exactly when its
// half-resolution pixel does). MODE 2 also stages the work-group's weight rows once, each
// row padded by 16 bytes so that the 16 rows of an A operand fall in distinct banks; the K
// loop then waits on no global load.
constexpr int kHalo = 18, kHaloUp = 10;
constexpr int kPix = kHalo * kHalo, kPixUp = kHaloUp * kHaloUp;
__host__ __device__ inline uint32_t conv_tile_bytes(uint32_t ca, uint32_t cb, uint32_t a_up) {
    return a_up ? (kPix * (cb / 8u) + kPixUp * (ca / 8u)) * 16u : kPix * ((ca + cb) / 8u) * 16u;
}
__host__ __device__ inline uint32_t conv_weight_bytes(uint32_t mt, uint32_t kpad) {
    return mt * 16u * (kpad / 8u + 1u) * 16u;
}

template <int MT, int MODE>
__global__ __launch_bounds__(256) void k_conv(ConvK K) {
    constexpr bool STAGE = MODE >= 1;
    extern __shared__ uint4 tile[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int img = blockIdx.y / K.tiles_y, ty = blockIdx.y - img * K.tiles_y;
    const int W = int(K.W), H = int(K.H);
    const int x = blockIdx.x * 16 + (lane & 15);
    const int y0 = ty * 16 + wv * 4;
    const int co0 = blockIdx.z * (MT * 16);
    const int kg = (lane >> 4) * 8;
    const int cin = int(K.cin), ca = int(K.ca), cb = int(K.cb);
    const size_t px0 = size_t(img) * H * W;
    const size_t pxa0 = size_t(img) * (H >> 1) * (W >> 1);
    f32x4 acc[MT][4];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wrow = K.w + size_t(co0 + (lane & 15)) * K.kpad + kg;
    // full-resolution region: channels [cfull0, cin) of each of the 18 x 18 pixels
    const int cfull0 = K.a_up ? ca : 0;
    const int pstride = (cin - cfull0) / 8;
    const int ustride = ca / 8;                          // half-resolution region (a_up)
    const int ubase = kHalo * kHalo * pstride;
    const int bx0 = int(blockIdx.x) * 16, by0 = ty * 16;
    const int ux0 = (bx0 - 1) >> 1, uy0 = (by0 - 1) >> 1;
    const int wbase = ubase + (K.a_up ? kPixUp * ustride : 0);  // MODE 2: weight rows
    const int kq = int(K.kpad) >> 3, wst = kq + 1;
    if (STAGE) {
        for (int u = int(threadIdx.x); u < kPix * pstride; u += 256) {
            const int pix = u / pstride, g = u - pix * pstride;
            const int r = pix / kHalo, c = pix - r * kHalo;
            const int sy = by0 - 1 + r, sx = bx0 - 1 + c, c8 = cfull0 + g * 8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
                const uint16_t* src = c8 < ca ? K.a + (px0 + size_t(sy) * W + sx) * ca + c8
                                              : K.b + (px0 + size_t(sy) * W + sx) * cb + (c8 - ca);
                v = *reinterpret_cast<const uint4*>(src);
            }
            tile[g * kPix + pix] = v;
        }
        if (K.a_up) {
            for (int u = int(threadIdx.x); u < kPixUp * ustride; u += 256) {
                const int pix = u / ustride, g = u - pix * ustride;
                const int r = pix / kHaloUp, c = pix - r * kHaloUp;
                const int uy = uy0 + r, ux = ux0 + c;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (ux >= 0 && ux < (W >> 1) && uy >= 0 && uy < (H >> 1))
                    v = *reinterpret_cast<const uint4*>(K.a + (pxa0 + size_t(uy) * (W >> 1) + ux) * ca + g * 8);
                tile[ubase + g * kPixUp + pix] = v;
            }
        }
        if (MODE == 2) {
            for (int u = int(threadIdx.x); u < MT * 16 * kq; u += 256) {
                const int row = u / kq, col = u - row * kq;
                tile[wbase + row * wst + col] =
                    *reinterpret_cast<const uint4*>(K.w + size_t(co0 + row) * K.kpad + size_t(col) * 8);
            }
        }
        __syncthreads();
    }
    int tap = kg / cin, ci = kg - tap * cin;
    for (uint32_t kc = 0; kc < K.kpad; kc += 32) {
        bf16x8 af[MT];
        if (MODE == 2) {
#pragma unroll
            for (int i = 0; i < MT; ++i)
                af[i] = __builtin_bit_cast(bf16x8, tile[wbase + (i * 16 + (lane & 15)) * wst + int((kc + uint32_t(kg)) >> 3)]);
        } else {
#pragma unroll
            for (int i = 0; i < MT; ++i) af[i] = ld8(wrow + size_t(i * 16) * K.kpad + kc);
        }
        bf16x8 bfr[4];
        const uint32_t kk = kc + uint32_t(kg);
        if (kk < K.ktap) {
            const int ty3 = tap / 3;
            const int dy = ty3 - 1, dx = tap - ty3 * 3 - 1;
            if (STAGE) {
                const int lx = (lane & 15) + dx + 1;
                if (ci < cfull0) {  // upsampled input: its half-resolution pixel
                    const int ucx = ((bx0 - 1 + lx) >> 1) - ux0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int ucy = ((by0 + wv * 4 + j + dy) >> 1) - uy0;
                        bfr[j] = __builtin_bit_cast(bf16x8, tile[ubase + (ci >> 3) * kPixUp + ucy * kHaloUp + ucx]);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        bfr[j] = __builtin_bit_cast(
                            bf16x8, tile[((ci - cfull0) >> 3) * kPix + (wv * 4 + j + dy + 1) * kHalo + lx]);
                }
            } else {
                const int sx = x + dx;
                const bool xin = sx >= 0 && sx < W;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int sy = y0 + j + dy;
                    bfr[j] = zero8();
                    if (xin && sy >= 0 && sy < H) {
                        const uint16_t* src;
                        if (ci < ca)
                            src = K.a_up ? K.a + (pxa0 + size_t(sy >> 1) * (W >> 1) + (sx >> 1)) * ca + ci
                                         : K.a + (px0 + size_t(sy) * W + sx) * ca + ci;
                        else
                            src = K.b + (px0 + size_t(sy) * W + sx) * cb + (ci - ca);
                        bfr[j] = ld8(src);
                    }
                }
            }
        } else {
            const bf16x8 v = (kk == K.ktap && K.bias) ? one8() : zero8();
#pragma unroll
            for (int j = 0; j < 4; ++j) bfr[j] = v;
        }
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        ci += 32;
        while (ci >= cin) {
            ci -= cin;
            ++tap;
        }
    }
    const int split = int(K.split), cout = int(K.cout);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = y0 + j;
        if (x >= W || y >= H) continue;
        const size_t px = px0 + size_t(y) * W + x;
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            const int co = co0 + i * 16 + (lane >> 4) * 4;
            float v0 = acc[i][j][0], v1 = acc[i][j][1], v2 = acc[i][j][2], v3 = acc[i][j][3];
            if (K.relu) {
                v0 = fmaxf(v0, 0.f);
                v1 = fmaxf(v1, 0.f);
                v2 = fmaxf(v2, 0.f);
                v3 = fmaxf(v3, 0.f);
            }
            if (K.outf) {
                if (co == 0) *reinterpret_cast<float4*>(K.outf + px * 4) = make_float4(v0, v1, v2, v3);
            } else if (co < split) {
                st4(K.out + px * split + co, v0, v1, v2, v3);
            } else {
                st4(K.out2 + px * (cout - split) + (co - split), v0, v1, v2, v3);
            }
        }
    }
}

// ---------------------------------------------------------------- forward convolution, fixed shape
// k_conv with the layer's channel counts as template constants (the layers of the shipped
// 32-channel network): the K loop is unrolled, so every LDS operand address is a per-lane
// base plus an immediate and the weight loads are a per-lane pointer plus an immediate; the
// runtime-shape kernel spends most of its issue slots on that index arithmetic. Same
// operands, same MFMA sequence and the same RNE rounding to BF16 (the hardware conversion),
// so the outputs equal k_conv's bit for bit. POOL: the epilogue also writes the 2 x 2
// average of the rounded outputs at half resolution, summed in k_pool's order.
struct ConvS {
    const uint16_t* a;   // first input (half resolution when UP)
    const uint16_t* b;   // skip input
    const uint16_t* w;   // [COUT][KPAD]
    uint16_t* out;       // NHWC, COUT channels
    uint16_t* pool;      // NHWC at half resolution (POOL)
    float* outf;         // float4 per pixel: channels 0..3 (OUT 1)
    // IN 1: FP32 features (vw x vh x 10); IN 2: the frame's targets (vw x vh), as k_frame_input
    const float* feat;
    const ushort4 *fd, *fs, *fa, *fn;
    const float4* fh;
    // OUT 2: linear radiance, 3 floats per pixel (vw x vh); OUT 3: the display image
    float* orgb;
    ushort4* odisp;
    uint32_t vw, vh;     // valid (unpadded) image within the W x H tensor
    float paper_white;
    uint32_t hdr, linear;
    uint32_t W, H, tiles_y, relu;
    uint32_t tiles;      // images x tile rows x tile columns (set by the launcher)
};

// R: 16-pixel rows per wave (the work-group's tile is 16 x 4R pixels)
template <int CA, int CB, bool UP, int R = 4>
struct ConvShape {
    static constexpr int CIN = CA + CB;
    static constexpr int KTAP = 9 * CIN;
    static constexpr int KPAD = (KTAP + 1 + 31) / 32 * 32;
    static constexpr int NSTEP = KPAD / 32;
    static constexpr int NPF = (UP ? CB : CIN) / 8;  // full-resolution planes
    static constexpr int NPU = UP ? CA / 8 : 0;      // half-resolution planes
    static constexpr int PIX = kHalo * (4 * R + 2);  // full-resolution plane: 18 x (4R + 2)
    static constexpr int PIXU = kHaloUp * (2 * R + 2);  // half-resolution plane: 10 x (2R + 2)
    static constexpr int UBASE = PIX * NPF;
    static constexpr int UNITS = UBASE + PIXU * NPU;
    // operand of k-lane-group starting at k, row j of the wave: kind 0 full-resolution plane,
    // 1 half-resolution plane, 2 bias one, 3 zero; off = the address's constant part in
    // 16-byte units; dxi = dx + 1 (selects the lane's half-resolution column)
    struct Op {
        int kind, off, dxi;
    };
    static constexpr Op op(int k, int j) {
        if (k >= KTAP) return Op{k == KTAP ? 2 : 3, 0, 0};
        const int tap = k / CIN, ci = k - (k / CIN) * CIN;
        const int dy = tap / 3 - 1, dx = tap - (tap / 3) * 3 - 1;
        if (UP && ci < CA) {
            // half-resolution row: floor((R wv + j + dy) / 2) + 1 = (R / 2) wv + floor((j + dy) / 2) + 1
            const int t = j + dy;
            const int fl = t >= 0 ? t / 2 : -1;
            return Op{1, UBASE + (ci / 8) * PIXU + (fl + 1) * kHaloUp, dx + 1};
        }
        const int plane = (UP ? ci - CA : ci) / 8;
        return Op{0, plane * PIX + (j + dy + 1) * kHalo + (dx + 1), 0};
    }
};

// Persistent: work-group b owns the contiguous run of tiles [b T / G, (b + 1) T / G) of the
// T = images x tile rows x tile columns 16 x 16 tiles (G = gridDim.x), so one work-group's
// consecutive tiles are neighbours and share their halo columns in its XCD's L2. WL: the
// weight rows are copied to LDS once per work-group (16 bytes of padding per row, so that
// the 16 rows of an A operand fall in distinct banks). DB: two input buffers; the next
// tile's global loads are issued into registers before the current tile's K loop and
// written to the other buffer after its epilogue, so their latency overlaps the MFMAs and
// the stores. LDS: [weights][buffer 0][buffer 1], a buffer holding an input tile and, after
// the K loop, the output tile.
// IN: 0 the BF16 tensor a, 1 / 2 the network input computed from FP32 features / the frame's
// targets (first layer); OUT: 0 BF16 tensor, 1 float4 per pixel, 2 linear radiance (k_unpack),
// 3 the display image (k_unpack_display) (last layer)
template <int CA, int CB, bool UP, int COUT, bool POOL, int OUT, bool WL, int IN, bool DB>
struct ConvP {
    using S = ConvShape<CA, CB, UP>;
    static constexpr int MT = COUT / 16;
    static constexpr int KQ = S::KPAD / 8, WST = KQ + 1;
    static constexpr int WUNITS = WL ? COUT * WST : 0;
    static constexpr int OUNITS = OUT != 0 ? 256 : 256 * COUT / 8;
    static constexpr int BUFU = S::UNITS > OUNITS ? S::UNITS : OUNITS;
    static constexpr int LDSU = WUNITS + (DB ? 2 : 1) * BUFU;
    // registers per thread holding the next tile: 16-byte units, or (IN != 0) pixels
    static constexpr int NPRE = IN != 0 ? (kPix + 255) / 256 : (S::UNITS + 255) / 256;
    struct Raw {  // a computed-input pixel's sources
        float v[10];
        bool in;
    };
};

template <int CA, int CB, bool UP, int COUT, bool POOL, int OUT, bool WL, int IN, bool DB>
__global__ __launch_bounds__(256) void k_convs(ConvS K) {
    using P = ConvP<CA, CB, UP, COUT, POOL, OUT, WL, IN, DB>;
    using S = typename P::S;
    constexpr int MT = P::MT, KQ = P::KQ, WST = P::WST;
    constexpr bool OUTF = OUT != 0;
    extern __shared__ uint4 lds[];
    uint4* const wts = lds;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gi = lane >> 4, x = lane & 15;
    const int W = int(K.W), H = int(K.H);
    const int ntx = (W + 15) >> 4, nty = int(K.tiles_y);
    const int T = int(K.tiles);
    const int t_begin = int((int64_t(blockIdx.x) * T) / gridDim.x);
    const int t_end = int((int64_t(blockIdx.x + 1) * T) / gridDim.x);
    if (t_begin >= t_end) return;
    typedef float f2v __attribute__((ext_vector_type(2)));
    typedef __bf16 b2v __attribute__((ext_vector_type(2)));

    // ---- next-tile loads into registers (zero outside the image)
    uint4 pre[DB && IN == 0 ? P::NPRE : 1];           // DB: the next tile's units
    typename P::Raw raw[IN != 0 && DB ? P::NPRE : 1];  // DB, computed input: sources, converted when stored
    // a computed-input pixel's 16 channels as two 16-byte units of BF16
    auto pack_input = [&](const typename P::Raw& r, uint4& lo, uint4& hi) {
        uint32_t q[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        if (r.in) {
            float f[16];
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] = log1pf(fmaxf(r.v[c], 0.f));
#pragma unroll
            for (int c = 3; c < 9; ++c) f[c] = r.v[c];
            f[9] = IN == 2 ? log1pf(fminf(fmaxf(r.v[9], 0.f), 1e4f)) * 0.1f : r.v[9];
#pragma unroll
            for (int c = 10; c < 16; ++c) f[c] = 0.f;
#pragma unroll
            for (int e = 0; e < 8; ++e)
                q[e] = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){f[2 * e], f[2 * e + 1]}, b2v));
        }
        lo = make_uint4(q[0], q[1], q[2], q[3]);
        hi = make_uint4(q[4], q[5], q[6], q[7]);
    };
    auto tile_xy = [&](int t, int& img, int& bx0, int& by0) {
        img = t / (ntx * nty);
        const int r = t - img * ntx * nty;
        by0 = (r / ntx) * 16;
        bx0 = (r - (r / ntx) * ntx) * 16;
    };
    auto load_tile = [&](int t) {
        int img, bx0, by0;
        tile_xy(t, img, bx0, by0);
        const size_t px0 = size_t(img) * H * W;
        if constexpr (IN != 0) {
#pragma unroll
            for (int k = 0; k < P::NPRE; ++k) {
                const int pix = int(threadIdx.x) + 256 * k;
                const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
                const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
                typename P::Raw& rk = raw[DB ? k : 0];
                rk.in = pix < kPix && sx >= 0 && sx < W && sy >= 0 && sy < H && sx < int(K.vw) && sy < int(K.vh);
                if (rk.in) {
                    const size_t i = size_t(sy) * K.vw + sx;
                    if (IN == 1) {
                        const float2* f2 = reinterpret_cast<const float2*>(K.feat + i * 10);  // 8-byte aligned
#pragma unroll
                        for (int e = 0; e < 5; ++e) {
                            const float2 q = f2[e];
                            rk.v[2 * e] = q.x;
                            rk.v[2 * e + 1] = q.y;
                        }
                    } else {
                        const float4 d = from_h4(K.fd[i]), sp = from_h4(K.fs[i]), al = from_h4(K.fa[i]),
                                     nm = from_h4(K.fn[i]);
                        // as k_frame_input; log1p of the depth at write time
                        rk.v[0] = d.x * al.x + sp.x;
                        rk.v[1] = d.y * al.y + sp.y;
                        rk.v[2] = d.z * al.z + sp.z;
                        rk.v[3] = al.x;
                        rk.v[4] = al.y;
                        rk.v[5] = al.z;
                        rk.v[6] = nm.x;
                        rk.v[7] = nm.y;
                        rk.v[8] = nm.z;
                        rk.v[9] = K.fh[i].x;
                    }
                }
            }
        } else {
            const int ux0 = (bx0 >> 1) - 1, uy0 = (by0 >> 1) - 1;
            const int Wh = W >> 1, Hh = H >> 1;
            const size_t pxa0 = size_t(img) * Hh * Wh;
#pragma unroll
            for (int k = 0; k < P::NPRE; ++k) {
                const int u = int(threadIdx.x) + 256 * k;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (u < kPix * S::NPF) {
                    const int pix = u / S::NPF, g = u - (u / S::NPF) * S::NPF;
                    const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
                    const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
                    if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
                        const size_t p = px0 + size_t(sy) * W + sx;
                        const int c8 = g * 8;
                        const uint16_t* src =
                            UP ? K.b + p * CB + c8 : (c8 < CA ? K.a + p * CA + c8 : K.b + p * CB + (c8 - CA));
                        v = *reinterpret_cast<const uint4*>(src);
                    }
                } else if (UP && u < S::UNITS) {
                    constexpr int NPU = S::NPU > 0 ? S::NPU : 1;
                    const int q = u - kPix * S::NPF;
                    const int pix = q / NPU, g = q - (q / NPU) * NPU;
                    const int r = pix / kHaloUp, c = pix - (pix / kHaloUp) * kHaloUp;
                    const int uy = uy0 + r, ux = ux0 + c;
                    if (ux >= 0 && ux < Wh && uy >= 0 && uy < Hh)
                        v = *reinterpret_cast<const uint4*>(K.a + (pxa0 + size_t(uy) * Wh + ux) * CA + g * 8);
                }
                pre[DB ? k : 0] = v;
            }
        }
    };
    // ---- the loaded tile into an LDS buffer (planes of 16-byte units per 8-channel group)
    auto store_tile = [&](uint4* buf) {
        if constexpr (IN != 0) {
#pragma unroll
            for (int k = 0; k < P::NPRE; ++k) {
                const int pix = int(threadIdx.x) + 256 * k;
                if (pix < kPix) {
                    uint4 lo, hi;
                    pack_input(raw[DB ? k : 0], lo, hi);
                    buf[pix] = lo;
                    buf[kPix + pix] = hi;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < P::NPRE; ++k) {
                const int u = int(threadIdx.x) + 256 * k;
                if (u < kPix * S::NPF) {
                    const int pix = u / S::NPF, g = u - (u / S::NPF) * S::NPF;
                    buf[g * kPix + pix] = pre[DB ? k : 0];
                } else if (UP && u < S::UNITS) {
                    constexpr int NPU = S::NPU > 0 ? S::NPU : 1;
                    const int q = u - kPix * S::NPF;
                    const int pix = q / NPU, g = q - (q / NPU) * NPU;
                    buf[S::UBASE + g * kPixUp + pix] = pre[DB ? k : 0];
                }
            }
        }
    };

    // ---- !DB: a tile straight into an LDS buffer, unit by unit
    auto stage_direct = [&](int t, uint4* buf) {
        int img, bx0, by0;
        tile_xy(t, img, bx0, by0);
        const size_t px0 = size_t(img) * H * W;
        if constexpr (IN != 0) {
            for (int pix = int(threadIdx.x); pix < kPix; pix += 256) {
                const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
                const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
                typename P::Raw rk;
                rk.in = sx >= 0 && sx < W && sy >= 0 && sy < H && sx < int(K.vw) && sy < int(K.vh);
                if (rk.in) {
                    const size_t i = size_t(sy) * K.vw + sx;
                    if (IN == 1) {
                        const float2* f2 = reinterpret_cast<const float2*>(K.feat + i * 10);
#pragma unroll
                        for (int e = 0; e < 5; ++e) {
                            const float2 q = f2[e];
                            rk.v[2 * e] = q.x;
                            rk.v[2 * e + 1] = q.y;
                        }
                    } else {
                        const float4 d = from_h4(K.fd[i]), sp = from_h4(K.fs[i]), al = from_h4(K.fa[i]),
                                     nm = from_h4(K.fn[i]);
                        rk.v[0] = d.x * al.x + sp.x;
                        rk.v[1] = d.y * al.y + sp.y;
                        rk.v[2] = d.z * al.z + sp.z;
                        rk.v[3] = al.x;
                        rk.v[4] = al.y;
                        rk.v[5] = al.z;
                        rk.v[6] = nm.x;
                        rk.v[7] = nm.y;
                        rk.v[8] = nm.z;
                        rk.v[9] = K.fh[i].x;
                    }
                }
                uint4 lo, hi;
                pack_input(rk, lo, hi);
                buf[pix] = lo;
                buf[kPix + pix] = hi;
            }
        } else {
            for (int u = int(threadIdx.x); u < kPix * S::NPF; u += 256) {
                const int pix = u / S::NPF, g = u - (u / S::NPF) * S::NPF;
                const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
                const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
                    const size_t p = px0 + size_t(sy) * W + sx;
                    const int c8 = g * 8;
                    const uint16_t* src = UP ? K.b + p * CB + c8 : (c8 < CA ? K.a + p * CA + c8 : K.b + p * CB + (c8 - CA));
                    v = *reinterpret_cast<const uint4*>(src);
                }
                buf[g * kPix + pix] = v;
            }
            if constexpr (UP) {
                const int Wh = W >> 1, Hh = H >> 1;
                const int ux0 = (bx0 >> 1) - 1, uy0 = (by0 >> 1) - 1;
                const size_t pxa0 = size_t(img) * Hh * Wh;
                constexpr int NPU = S::NPU > 0 ? S::NPU : 1;
                for (int u = int(threadIdx.x); u < kPixUp * S::NPU; u += 256) {
                    const int pix = u / NPU, g = u - (u / NPU) * NPU;
                    const int r = pix / kHaloUp, c = pix - (pix / kHaloUp) * kHaloUp;
                    const int uy = uy0 + r, ux = ux0 + c;
                    uint4 v = make_uint4(0u, 0u, 0u, 0u);
                    if (ux >= 0 && ux < Wh && uy >= 0 && uy < Hh)
                        v = *reinterpret_cast<const uint4*>(K.a + (pxa0 + size_t(uy) * Wh + ux) * CA + g * 8);
                    buf[S::UBASE + g * kPixUp + pix] = v;
                }
            }
        }
    };

    // ---- prologue: weights and the first tile
    if constexpr (WL) {
        for (int u = int(threadIdx.x); u < COUT * KQ; u += 256) {
            const int row = u / KQ, col = u - (u / KQ) * KQ;
            wts[row * WST + col] = *reinterpret_cast<const uint4*>(K.w + size_t(row) * S::KPAD + col * 8);
        }
    }
    uint4* buf_a = lds + P::WUNITS;
    uint4* buf_b = DB ? buf_a + P::BUFU : buf_a;
    if (DB) {
        load_tile(t_begin);
        store_tile(buf_a);
    } else {
        stage_direct(t_begin, buf_a);
    }
    __syncthreads();

    // ---- per-lane operand bases
    const int fb0 = wv * 4 * kHalo + x;  // full-resolution pixel of row 0
    // half-resolution column of the pixel for dx = -1, 0, 1: floor((x + dx) / 2) + 1, plus
    // the wave's first half-resolution row (2 wv rows of 10)
    const int u0 = 2 * wv * kHaloUp + ((x + 1) >> 1), u1 = 2 * wv * kHaloUp + ((x + 2) >> 1),
              u2 = 2 * wv * kHaloUp + ((x + 3) >> 1);
    auto ub0 = [&](int d) { return d == 0 ? u0 : (d == 1 ? u1 : u2); };
    const int fbg = fb0 + gi * kPix;
    auto ubg = [&](int d) { return ub0(d) + gi * kPixUp; };
    const uint16_t* wl = K.w + size_t(x) * S::KPAD + gi * 8;

    for (int t = t_begin; t < t_end; ++t) {
        uint4* const tile = buf_a;
        const bool more = t + 1 < t_end;
        if (DB && more) load_tile(t + 1);
        f32x4 acc[MT][4];
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S::NSTEP; ++s) {
            bf16x8 af[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i)
                af[i] = WL ? __builtin_bit_cast(bf16x8, wts[(i * 16 + x) * WST + s * 4 + gi])
                           : ld8(wl + i * 16 * S::KPAD + s * 32);
            bf16x8 bfr[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const typename S::Op q0 = S::op(32 * s, j), q1 = S::op(32 * s + 8, j);
                const typename S::Op q2 = S::op(32 * s + 16, j), q3 = S::op(32 * s + 24, j);
                const int psz = q0.kind == 0 ? kPix : kPixUp;
                const bool uniform = q0.kind <= 1 && q1.kind == q0.kind && q2.kind == q0.kind && q3.kind == q0.kind &&
                                     q1.dxi == q0.dxi && q2.dxi == q0.dxi && q3.dxi == q0.dxi &&
                                     q1.off == q0.off + psz && q2.off == q0.off + 2 * psz && q3.off == q0.off + 3 * psz;
                if (uniform) {
                    const int off = q0.off + (q0.kind == 0 ? fbg : ubg(q0.dxi));
                    bfr[j] = __builtin_bit_cast(bf16x8, tile[off]);
                } else {
                    auto lo = [&](const typename S::Op& q) {
                        return q.kind == 0 ? q.off + fb0 : (q.kind == 1 ? q.off + ub0(q.dxi) : 0);
                    };
                    const int off = gi == 0 ? lo(q0) : gi == 1 ? lo(q1) : gi == 2 ? lo(q2) : lo(q3);
                    const int kind = gi == 0 ? q0.kind : gi == 1 ? q1.kind : gi == 2 ? q2.kind : q3.kind;
                    bf16x8 v = __builtin_bit_cast(bf16x8, tile[off]);
                    if (q0.kind >= 2 || q1.kind >= 2 || q2.kind >= 2 || q3.kind >= 2) {
                        if (kind == 2) v = one8();
                        else if (kind == 3) v = zero8();
                    }
                    bfr[j] = v;
                }
            }
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        // ---- epilogue: the output tile goes through the buffer ([pixel][channel]) so that
        // the global stores are whole 16-byte units of consecutive pixels (a lane's own 4
        // channels of one pixel per store left each cache line half written per instruction)
        int img, bx0, by0;
        tile_xy(t, img, bx0, by0);
        const size_t px0 = size_t(img) * H * W;
        __syncthreads();  // every wave is done reading the input tile
        uint16_t* otile = reinterpret_cast<uint16_t*>(tile);
        float4* ftile = reinterpret_cast<float4*>(tile);
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            const int co = i * 16 + gi * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v0 = acc[i][j][0], v1 = acc[i][j][1], v2 = acc[i][j][2], v3 = acc[i][j][3];
                if (K.relu) {
                    v0 = fmaxf(v0, 0.f);
                    v1 = fmaxf(v1, 0.f);
                    v2 = fmaxf(v2, 0.f);
                    v3 = fmaxf(v3, 0.f);
                }
                const int p = (wv * 4 + j) * 16 + x;
                if (OUTF) {
                    if (co == 0) ftile[p] = make_float4(v0, v1, v2, v3);
                } else {
                    const b2v p01 = __builtin_convertvector((f2v){v0, v1}, b2v);
                    const b2v p23 = __builtin_convertvector((f2v){v2, v3}, b2v);
                    *reinterpret_cast<uint2*>(otile + p * COUT + co) =
                        make_uint2(__builtin_bit_cast(uint32_t, p01), __builtin_bit_cast(uint32_t, p23));
                }
            }
        }
        __syncthreads();
        constexpr int CU = OUTF ? 1 : COUT / 8;  // 16-byte units per pixel
        for (int c = int(threadIdx.x); c < 256 * CU; c += 256) {
            const int p = c / CU, g = c - (c / CU) * CU;
            const int gy = by0 + (p >> 4), gx = bx0 + (p & 15);
            if (gx < W && gy < H) {
                const size_t px = px0 + size_t(gy) * W + gx;
                if (OUT == 0) {
                    *reinterpret_cast<uint4*>(K.out + px * COUT + g * 8) = tile[p * CU + g];
                } else if (OUT == 1) {
                    *reinterpret_cast<uint4*>(K.outf + px * 4) = tile[p];
                } else if (gx < int(K.vw) && gy < int(K.vh)) {
                    const float4 z = ftile[p];
                    const float r = fmaxf(expm1f(z.x), 0.f), gg = fmaxf(expm1f(z.y), 0.f), bb = fmaxf(expm1f(z.z), 0.f);
                    const size_t tt = size_t(gy) * K.vw + gx;
                    if (OUT == 2) {
                        K.orgb[3 * tt] = r;
                        K.orgb[3 * tt + 1] = gg;
                        K.orgb[3 * tt + 2] = bb;
                    } else {
                        K.odisp[tt] = K.linear ? to_h4(r, gg, bb, 1.f)
                                               : to_h4(display_encode(r, K.paper_white, K.hdr != 0u),
                                                       display_encode(gg, K.paper_white, K.hdr != 0u),
                                                       display_encode(bb, K.paper_white, K.hdr != 0u), 1.f);
                    }
                }
            }
        }
        if (POOL) {
            // 2 x 2 average of the rounded outputs in k_pool's order: (y, x), (y, x + 1),
            // (y + 1, x), (y + 1, x + 1)
            const int Wh = W >> 1, Hh = H >> 1;
            for (int c = int(threadIdx.x); c < 64 * CU; c += 256) {
                const int q = c / CU, g = c - (c / CU) * CU;
                const int qy = q >> 3, qx = q & 7;
                const int p00 = (2 * qy) * 16 + 2 * qx;
                const uint4 t4[4] = {tile[p00 * CU + g], tile[(p00 + 1) * CU + g], tile[(p00 + 16) * CU + g],
                                     tile[(p00 + 17) * CU + g]};
                float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const uint32_t w4[4] = {t4[d].x, t4[d].y, t4[d].z, t4[d].w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        o[2 * e] += __uint_as_float(w4[e] << 16);
                        o[2 * e + 1] += __uint_as_float(w4[e] & 0xffff0000u);
                    }
                }
                uint32_t r4[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    r4[e] = __builtin_bit_cast(uint32_t,
                                               __builtin_convertvector((f2v){0.25f * o[2 * e], 0.25f * o[2 * e + 1]}, b2v));
                const int py = (by0 >> 1) + qy, pxh = (bx0 >> 1) + qx;
                if (pxh < Wh && py < Hh) {
                    const size_t pp = size_t(img) * Hh * Wh + size_t(py) * Wh + pxh;
                    *reinterpret_cast<uint4*>(K.pool + pp * COUT + g * 8) = make_uint4(r4[0], r4[1], r4[2], r4[3]);
                }
            }
        }
        if (!more) break;
        if (DB) {
            // the other buffer is free (its last use ended before this tile's K loop)
            store_tile(buf_b);
        } else {
            __syncthreads();  // the output tile's reads are done before the buffer is refilled
            stage_direct(t + 1, buf_a);
        }
        __syncthreads();
        if (DB) {
            uint4* const sw = buf_a;
            buf_a = buf_b;
            buf_b = sw;
        }
    }
}

// the network's 16 input channels of pixel (x, y) of the valid image, computed from FP32
// features (IN 1) or the frame's targets (IN 2) as k_frame_input and k_pack do:
// log(1 + max(radiance, 0)), albedo, normal, depth feature, zeros
template <int IN>
__device__ inline void conv_input(const ConvS& K, int x, int y, float (&f)[16]) {
    const size_t i = size_t(y) * K.vw + x;
    float v[10];
    if (IN == 1) {
        const float2* f2 = reinterpret_cast<const float2*>(K.feat + i * 10);  // 8-byte aligned
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float2 t = f2[c];
            v[2 * c] = t.x;
            v[2 * c + 1] = t.y;
        }
    } else {
        const float4 d = from_h4(K.fd[i]), s = from_h4(K.fs[i]), a = from_h4(K.fa[i]), nm = from_h4(K.fn[i]);
        v[0] = d.x * a.x + s.x;
        v[1] = d.y * a.y + s.y;
        v[2] = d.z * a.z + s.z;
        v[3] = a.x;
        v[4] = a.y;
        v[5] = a.z;
        v[6] = nm.x;
        v[7] = nm.y;
        v[8] = nm.z;
        v[9] = log1pf(fminf(fmaxf(K.fh[i].x, 0.f), 1e4f)) * 0.1f;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = log1pf(fmaxf(v[c], 0.f));
#pragma unroll
    for (int c = 3; c < 10; ++c) f[c] = v[c];
#pragma unroll
    for (int c = 10; c < 16; ++c) f[c] = 0.f;
}

// One tile per work-group (the layers where persistence does not pay, see launch_convs_t):
// the same K loop and epilogue as k_convs, the input staged straight into LDS.
template <int CA, int CB, bool UP, int COUT, bool POOL, bool OUTF, bool WL, int IN, int R>
__global__ __launch_bounds__(256) void k_convs1(ConvS K) {
    using S = ConvShape<CA, CB, UP, R>;
    constexpr int MT = COUT / 16;
    constexpr int KQ = S::KPAD / 8, WST = KQ + 1;  // WL: weight rows in LDS, 16 bytes of padding each
    extern __shared__ uint4 tile[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gi = lane >> 4, x = lane & 15;
    const int img = blockIdx.y / K.tiles_y, ty = blockIdx.y - img * K.tiles_y;
    const int W = int(K.W), H = int(K.H);
    const int bx0 = int(blockIdx.x) * 16, by0 = ty * 4 * R;
    const size_t px0 = size_t(img) * H * W;
    // ---- stage the input tile (planes of 16-byte units, zero outside the image)
    if constexpr (IN != 0) {
        // the network input of each pixel computed once, both of its planes written
        static_assert(CA == 16 && CB == 0 && !UP, "computed input: first layer only");
        typedef float f2i __attribute__((ext_vector_type(2)));
        typedef __bf16 b2i __attribute__((ext_vector_type(2)));
        for (int pix = int(threadIdx.x); pix < S::PIX; pix += 256) {
            const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
            const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
            uint32_t q[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
            if (sx >= 0 && sx < W && sy >= 0 && sy < H && sx < int(K.vw) && sy < int(K.vh)) {
                float f[16];
                conv_input<IN>(K, sx, sy, f);
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    q[e] = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2i){f[2 * e], f[2 * e + 1]}, b2i));
            }
            tile[pix] = make_uint4(q[0], q[1], q[2], q[3]);
            tile[S::PIX + pix] = make_uint4(q[4], q[5], q[6], q[7]);
        }
    } else {
        constexpr int NF = S::PIX * S::NPF;
        for (int u = int(threadIdx.x); u < NF; u += 256) {
            const int pix = u / S::NPF, g = u - (u / S::NPF) * S::NPF;
            const int r = pix / kHalo, c = pix - (pix / kHalo) * kHalo;
            const int sy = by0 - 1 + r, sx = bx0 - 1 + c;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
                const size_t p = px0 + size_t(sy) * W + sx;
                const int c8 = g * 8;
                const uint16_t* src = UP ? K.b + p * CB + c8 : (c8 < CA ? K.a + p * CA + c8 : K.b + p * CB + (c8 - CA));
                v = *reinterpret_cast<const uint4*>(src);
            }
            tile[g * S::PIX + pix] = v;
        }
        if constexpr (UP) {
            const int Wh = W >> 1, Hh = H >> 1;
            const int ux0 = (bx0 >> 1) - 1, uy0 = (by0 >> 1) - 1;
            const size_t pxa0 = size_t(img) * Hh * Wh;
            constexpr int NU = S::PIXU * S::NPU;
            for (int u = int(threadIdx.x); u < NU; u += 256) {
                const int pix = u / S::NPU, g = u - (u / S::NPU) * S::NPU;
                const int r = pix / kHaloUp, c = pix - (pix / kHaloUp) * kHaloUp;
                const int uy = uy0 + r, ux = ux0 + c;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (ux >= 0 && ux < Wh && uy >= 0 && uy < Hh)
                    v = *reinterpret_cast<const uint4*>(K.a + (pxa0 + size_t(uy) * Wh + ux) * CA + g * 8);
                tile[S::UBASE + g * S::PIXU + pix] = v;
            }
        }
    }
    if constexpr (WL) {
        for (int u = int(threadIdx.x); u < COUT * KQ; u += 256) {
            const int row = u / KQ, col = u - (u / KQ) * KQ;
            tile[S::UNITS + row * WST + col] = *reinterpret_cast<const uint4*>(K.w + size_t(row) * S::KPAD + col * 8);
        }
    }
    __syncthreads();
    // ---- per-lane operand bases
    const int fb0 = wv * R * kHalo + x;  // full-resolution pixel of row 0
    // half-resolution column of the pixel for dx = -1, 0, 1: floor((x + dx) / 2) + 1, plus
    // the wave's first half-resolution row (2 wv rows of 10)
    const int u0 = (R / 2) * wv * kHaloUp + ((x + 1) >> 1), u1 = (R / 2) * wv * kHaloUp + ((x + 2) >> 1),
              u2 = (R / 2) * wv * kHaloUp + ((x + 3) >> 1);
    auto ub0 = [&](int d) { return d == 0 ? u0 : (d == 1 ? u1 : u2); };
    const int fbg = fb0 + gi * S::PIX;
    auto ubg = [&](int d) { return ub0(d) + gi * S::PIXU; };
    const uint16_t* wl = K.w + size_t(x) * S::KPAD + gi * 8;
    f32x4 acc[MT][R];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < S::NSTEP; ++s) {
        bf16x8 af[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i)
            af[i] = WL ? __builtin_bit_cast(bf16x8, tile[S::UNITS + (i * 16 + x) * WST + s * 4 + gi])
                       : ld8(wl + i * 16 * S::KPAD + s * 32);
        bf16x8 bfr[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const typename S::Op q0 = S::op(32 * s, j), q1 = S::op(32 * s + 8, j);
            const typename S::Op q2 = S::op(32 * s + 16, j), q3 = S::op(32 * s + 24, j);
            const int psz = q0.kind == 0 ? S::PIX : S::PIXU;
            const bool uniform = q0.kind <= 1 && q1.kind == q0.kind && q2.kind == q0.kind && q3.kind == q0.kind &&
                                 q1.dxi == q0.dxi && q2.dxi == q0.dxi && q3.dxi == q0.dxi && q1.off == q0.off + psz &&
                                 q2.off == q0.off + 2 * psz && q3.off == q0.off + 3 * psz;
            if (uniform) {
                const int off = q0.off + (q0.kind == 0 ? fbg : ubg(q0.dxi));
                bfr[j] = __builtin_bit_cast(bf16x8, tile[off]);
            } else {
                auto lo = [&](const typename S::Op& q) {
                    return q.kind == 0 ? q.off + fb0 : (q.kind == 1 ? q.off + ub0(q.dxi) : 0);
                };
                const int off = gi == 0 ? lo(q0) : gi == 1 ? lo(q1) : gi == 2 ? lo(q2) : lo(q3);
                const int kind = gi == 0 ? q0.kind : gi == 1 ? q1.kind : gi == 2 ? q2.kind : q3.kind;
                bf16x8 v = __builtin_bit_cast(bf16x8, tile[off]);
                if (q0.kind >= 2 || q1.kind >= 2 || q2.kind >= 2 || q3.kind >= 2) {
                    if (kind == 2) v = one8();
                    else if (kind == 3) v = zero8();
                }
                bfr[j] = v;
            }
        }
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    // ---- epilogue: the output tile goes through LDS ([pixel][channel], the input tile's
    // space) so that the global stores are whole 16-byte units of consecutive pixels (a
    // lane's own 4 channels of one pixel per store left each cache line half written per
    // instruction and took most of the kernel's time)
    typedef float f2v __attribute__((ext_vector_type(2)));
    typedef __bf16 b2v __attribute__((ext_vector_type(2)));
    __syncthreads();  // every wave is done reading the input tile
    uint16_t* otile = reinterpret_cast<uint16_t*>(tile);
    float4* ftile = reinterpret_cast<float4*>(tile);
#pragma unroll
    for (int i = 0; i < MT; ++i) {
        const int co = i * 16 + gi * 4;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            float v0 = acc[i][j][0], v1 = acc[i][j][1], v2 = acc[i][j][2], v3 = acc[i][j][3];
            if (K.relu) {
                v0 = fmaxf(v0, 0.f);
                v1 = fmaxf(v1, 0.f);
                v2 = fmaxf(v2, 0.f);
                v3 = fmaxf(v3, 0.f);
            }
            const int p = (wv * R + j) * 16 + x;
            if (OUTF) {
                if (co == 0) ftile[p] = make_float4(v0, v1, v2, v3);
            } else {
                const b2v p01 = __builtin_convertvector((f2v){v0, v1}, b2v);
                const b2v p23 = __builtin_convertvector((f2v){v2, v3}, b2v);
                *reinterpret_cast<uint2*>(otile + p * COUT + co) =
                    make_uint2(__builtin_bit_cast(uint32_t, p01), __builtin_bit_cast(uint32_t, p23));
            }
        }
    }
    __syncthreads();
    constexpr int CU = OUTF ? 1 : COUT / 8;  // 16-byte units per pixel
    for (int c = int(threadIdx.x); c < 64 * R * CU; c += 256) {
        const int p = c / CU, g = c - (c / CU) * CU;
        const int gy = by0 + (p >> 4), gx = bx0 + (p & 15);
        if (gx < W && gy < H) {
            const size_t px = px0 + size_t(gy) * W + gx;
            if (OUTF) *reinterpret_cast<uint4*>(K.outf + px * 4) = tile[p];
            else *reinterpret_cast<uint4*>(K.out + px * COUT + g * 8) = tile[p * CU + g];
        }
    }
    if (POOL) {
        // 2 x 2 average of the rounded outputs in k_pool's order: (y, x), (y, x + 1),
        // (y + 1, x), (y + 1, x + 1)
        const int Wh = W >> 1, Hh = H >> 1;
        for (int c = int(threadIdx.x); c < 16 * R * CU; c += 256) {
            const int q = c / CU, g = c - (c / CU) * CU;
            const int qy = q >> 3, qx = q & 7;
            const int p00 = (2 * qy) * 16 + 2 * qx;
            const uint4 t4[4] = {tile[p00 * CU + g], tile[(p00 + 1) * CU + g], tile[(p00 + 16) * CU + g],
                                 tile[(p00 + 17) * CU + g]};
            float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const uint32_t w4[4] = {t4[d].x, t4[d].y, t4[d].z, t4[d].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    o[2 * e] += __uint_as_float(w4[e] << 16);
                    o[2 * e + 1] += __uint_as_float(w4[e] & 0xffff0000u);
                }
            }
            uint32_t r4[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                r4[e] = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){0.25f * o[2 * e], 0.25f * o[2 * e + 1]}, b2v));
            const int py = (by0 >> 1) + qy, pxh = (bx0 >> 1) + qx;
            if (pxh < Wh && py < Hh) {
                const size_t pp = size_t(img) * Hh * Wh + size_t(py) * Wh + pxh;
                *reinterpret_cast<uint4*>(K.pool + pp * COUT + g * 8) = make_uint4(r4[0], r4[1], r4[2], r4[3]);
            }
        }
    }
}

// ---------------------------------------------------------------- weight gradient
struct WgradK {
    const uint16_t* a;
    const uint16_t* b;
    const uint16_t* g;  // gradient at the convolution output (after the ReLU mask), [pixel][cout]
    float* dw;          // [cout][kpad], accumulated
    uint32_t ca, cb, cin, a_up;
    uint32_t W, H, rows, rows_per_block;
    uint32_t kpad, ktap, cout, bias;
};

// One work-group = 4 waves over rows_per_block image rows (the GEMM's reduction), MT*16
// output channels and NT*16 columns k. Each wave reduces its own rows and adds its tile
// to dw with atomics. Operands are gathered 2 bytes at a time (the reduction runs along
// pixels, which are not contiguous in NHWC for a fixed channel).
template <int MT, int NT>
__global__ __launch_bounds__(256) void k_wgrad(WgradK K) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int W = int(K.W), H = int(K.H);
    const int co0 = blockIdx.z * (MT * 16);
    const int kt0 = blockIdx.y * (NT * 16);
    const int kg = (lane >> 4) * 8;
    const int cin = int(K.cin), ca = int(K.ca), cb = int(K.cb), cout = int(K.cout);
    // column of this lane in each B tile: kind 0 zero, 1 data, 2 bias
    int kind[NT], dxj[NT], dyj[NT], cij[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int kcol = kt0 + j * 16 + (lane & 15);
        kind[j] = 0;
        dxj[j] = dyj[j] = cij[j] = 0;
        if (kcol < int(K.ktap)) {
            const int tap = kcol / cin;
            cij[j] = kcol - tap * cin;
            dyj[j] = tap / 3 - 1;
            dxj[j] = tap - (tap / 3) * 3 - 1;
            kind[j] = 1;
        } else if (kcol == int(K.ktap) && K.bias) {
            kind[j] = 2;
        }
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint32_t r_end = min(K.rows, (blockIdx.x + 1) * K.rows_per_block);
    for (uint32_t r = blockIdx.x * K.rows_per_block + wv; r < r_end; r += 4) {
        const int img = int(r / K.H), y = int(r - uint32_t(img) * K.H);
        const size_t px0 = size_t(img) * H * W;
        const size_t pxa0 = size_t(img) * (H >> 1) * (W >> 1);
        for (int x0 = 0; x0 < W; x0 += 32) {
            bf16x8 af[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) {
                const uint16_t* gp = K.g + (px0 + size_t(y) * W) * cout + co0 + i * 16 + (lane & 15);
                uint16_t v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int xx = x0 + kg + e;
                    v[e] = xx < W ? gp[size_t(xx) * cout] : uint16_t(0);
                }
                af[i] = pack8(v);
            }
            bf16x8 bfr[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                uint16_t v[8];
                const int sy = y + dyj[j];
                const bool yin = sy >= 0 && sy < H;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int xx = x0 + kg + e;
                    uint16_t val = 0;
                    if (xx < W) {
                        if (kind[j] == 2) {
                            val = 0x3f80u;
                        } else if (kind[j] == 1) {
                            const int sx = xx + dxj[j];
                            if (yin && sx >= 0 && sx < W) {
                                const int c = cij[j];
                                if (c < ca)
                                    val = K.a_up ? K.a[(pxa0 + size_t(sy >> 1) * (W >> 1) + (sx >> 1)) * ca + c]
                                                 : K.a[(px0 + size_t(sy) * W + sx) * ca + c];
                                else
                                    val = K.b[(px0 + size_t(sy) * W + sx) * cb + (c - ca)];
                            }
                        }
                    }
                    v[e] = val;
                }
                bfr[j] = pack8(v);
            }
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int kcol = kt0 + j * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int co = co0 + i * 16 + (lane >> 4) * 4 + e;
                const float v = acc[i][j][e];
                if (v != 0.f) unsafeAtomicAdd(K.dw + size_t(co) * K.kpad + kcol, v);
            }
        }
}

// ---------------------------------------------------------------- elementwise
__global__ void k_fill(float* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// 2x2 average pool, one thread per (output pixel, 8-channel group)
__global__ void k_pool(const uint16_t* s, uint16_t* p, uint32_t nimg, uint32_t H, uint32_t W, uint32_t c) {
    const uint32_t Ho = H / 2, Wo = W / 2, g8 = c / 8;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nimg) * Ho * Wo * g8) return;
    const uint32_t grp = uint32_t(t % g8);
    const size_t po = t / g8;
    const uint32_t x = uint32_t(po % Wo), y = uint32_t((po / Wo) % Ho), img = uint32_t(po / (size_t(Wo) * Ho));
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 4; ++d) {
        const size_t pi = (size_t(img) * H + 2 * y + (d >> 1)) * W + 2 * x + (d & 1);
        const uint4 v = *reinterpret_cast<const uint4*>(s + pi * c + grp * 8);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        for (int e = 0; e < 4; ++e) {
            acc[2 * e] += __uint_as_float(w4[e] << 16);
            acc[2 * e + 1] += __uint_as_float(w4[e] & 0xffff0000u);
        }
    }
    uint16_t* o = p + po * c + grp * 8;
    st4(o, 0.25f * acc[0], 0.25f * acc[1], 0.25f * acc[2], 0.25f * acc[3]);
    st4(o + 4, 0.25f * acc[4], 0.25f * acc[5], 0.25f * acc[6], 0.25f * acc[7]);
}

// gS (H x W) += gP (H/2 x W/2) / 4 broadcast to each 2x2 block
__global__ void k_pool_bwd(const uint16_t* gp, uint16_t* gs, uint32_t nimg, uint32_t H, uint32_t W, uint32_t c) {
    const uint32_t Ho = H / 2, Wo = W / 2, g8 = c / 8;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nimg) * Ho * Wo * g8) return;
    const uint32_t grp = uint32_t(t % g8);
    const size_t po = t / g8;
    const uint32_t x = uint32_t(po % Wo), y = uint32_t((po / Wo) % Ho), img = uint32_t(po / (size_t(Wo) * Ho));
    const uint4 v = *reinterpret_cast<const uint4*>(gp + po * c + grp * 8);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
    float q[8];
    for (int e = 0; e < 4; ++e) {
        q[2 * e] = 0.25f * __uint_as_float(w4[e] << 16);
        q[2 * e + 1] = 0.25f * __uint_as_float(w4[e] & 0xffff0000u);
    }
    for (int d = 0; d < 4; ++d) {
        const size_t pi = (size_t(img) * H + 2 * y + (d >> 1)) * W + 2 * x + (d & 1);
        uint16_t* o = gs + pi * c + grp * 8;
        const uint4 u = *reinterpret_cast<const uint4*>(o);
        const uint32_t s4[4] = {u.x, u.y, u.z, u.w};
        float r[8];
        for (int e = 0; e < 4; ++e) {
            r[2 * e] = __uint_as_float(s4[e] << 16) + q[2 * e];
            r[2 * e + 1] = __uint_as_float(s4[e] & 0xffff0000u) + q[2 * e + 1];
        }
        st4(o, r[0], r[1], r[2], r[3]);
        st4(o + 4, r[4], r[5], r[6], r[7]);
    }
}

// gA (H/2 x W/2) = sum of gUp over each 2x2 block (adjoint of nearest upsampling)
__global__ void k_upsum(const uint16_t* gu, uint16_t* ga, uint32_t nimg, uint32_t H, uint32_t W, uint32_t c) {
    const uint32_t Ho = H / 2, Wo = W / 2, g8 = c / 8;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nimg) * Ho * Wo * g8) return;
    const uint32_t grp = uint32_t(t % g8);
    const size_t po = t / g8;
    const uint32_t x = uint32_t(po % Wo), y = uint32_t((po / Wo) % Ho), img = uint32_t(po / (size_t(Wo) * Ho));
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 4; ++d) {
        const size_t pi = (size_t(img) * H + 2 * y + (d >> 1)) * W + 2 * x + (d & 1);
        const uint4 v = *reinterpret_cast<const uint4*>(gu + pi * c + grp * 8);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        for (int e = 0; e < 4; ++e) {
            acc[2 * e] += __uint_as_float(w4[e] << 16);
            acc[2 * e + 1] += __uint_as_float(w4[e] & 0xffff0000u);
        }
    }
    uint16_t* o = ga + po * c + grp * 8;
    st4(o, acc[0], acc[1], acc[2], acc[3]);
    st4(o + 4, acc[4], acc[5], acc[6], acc[7]);
}

// g = y > 0 ? g : 0, 8 bf16 per thread
__global__ void k_relu_mask(uint16_t* g, const uint16_t* y, size_t n8) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n8) return;
    uint4 gv = reinterpret_cast<uint4*>(g)[t];
    const uint4 yv = reinterpret_cast<const uint4*>(y)[t];
    auto m = [](uint32_t gw, uint32_t yw) {
        // bf16 > 0: sign bit clear and not +0
        const uint32_t lo = ((yw & 0x8000u) == 0u && (yw & 0x7fffu) != 0u) ? 0xffffu : 0u;
        const uint32_t hi = ((yw & 0x80000000u) == 0u && (yw & 0x7fff0000u) != 0u) ? 0xffff0000u : 0u;
        return gw & (lo | hi);
    };
    gv.x = m(gv.x, yv.x);
    gv.y = m(gv.y, yv.y);
    gv.z = m(gv.z, yv.z);
    gv.w = m(gv.w, yv.w);
    reinterpret_cast<uint4*>(g)[t] = gv;
}

// features (n x h x w x 10, FP32) -> X (n x hp x wp x 16, BF16): log(1+radiance), albedo,
// normal, depth feature, zeros; pixels outside h x w are zero
__global__ void k_pack(const float* in, uint16_t* X, uint32_t nimg, uint32_t h, uint32_t w, uint32_t hp, uint32_t wp) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nimg) * hp * wp) return;
    const uint32_t x = uint32_t(t % wp), y = uint32_t((t / wp) % hp), img = uint32_t(t / (size_t(wp) * hp));
    float f[16];
    for (int c = 0; c < 16; ++c) f[c] = 0.f;
    if (x < w && y < h) {
        const float* s = in + ((size_t(img) * h + y) * w + x) * kIn;
        for (int c = 0; c < 3; ++c) f[c] = log1pf(fmaxf(s[c], 0.f));
        for (int c = 3; c < int(kIn); ++c) f[c] = s[c];
    }
    uint16_t* o = X + t * kInPad;
    for (int c = 0; c < 16; c += 4) st4(o + c, f[c], f[c + 1], f[c + 2], f[c + 3]);
}

// network output (n x hp x wp float4, log radiance) -> n x h x w x 3 linear radiance
__global__ void k_unpack(const float* Y, float* out, uint32_t nimg, uint32_t h, uint32_t w, uint32_t hp, uint32_t wp) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nimg) * h * w) return;
    const uint32_t x = uint32_t(t % w), y = uint32_t((t / w) % h), img = uint32_t(t / (size_t(w) * h));
    const float4 z = reinterpret_cast<const float4*>(Y)[(size_t(img) * hp + y) * wp + x];
    out[3 * t] = fmaxf(expm1f(z.x), 0.f);
    out[3 * t + 1] = fmaxf(expm1f(z.y), 0.f);
    out[3 * t + 2] = fmaxf(expm1f(z.z), 0.f);
}

// network output -> tonemapped display image (GK_RT_DENOISED, RGBA16F), as the bilateral
// path writes it
__global__ void k_unpack_display(const float* Y, ushort4* out, uint32_t h, uint32_t w, uint32_t wp,
                                 float paper_white, uint32_t hdr, uint32_t linear) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= h * w) return;
    const uint32_t x = t % w, y = t / w;
    const float4 z = reinterpret_cast<const float4*>(Y)[size_t(y) * wp + x];
    const float r = fmaxf(expm1f(z.x), 0.f), g = fmaxf(expm1f(z.y), 0.f), b = fmaxf(expm1f(z.z), 0.f);
    out[t] = linear ? to_h4(r, g, b, 1.f)
                    : to_h4(display_encode(r, paper_white, hdr != 0u), display_encode(g, paper_white, hdr != 0u),
                            display_encode(b, paper_white, hdr != 0u), 1.f);
}

// loss = sum (z - log(1+t))^2 / (3 n h w); g = d loss / d z as BF16 (16 channels)
__global__ __launch_bounds__(256) void k_loss(const float* Y, const float* target, uint16_t* g, size_t npx,
                                              float inv_count, double* loss) {
    __shared__ float red[256];
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < npx) {
        const float4 z = reinterpret_cast<const float4*>(Y)[t];
        const float d0 = z.x - log1pf(fmaxf(target[3 * t], 0.f));
        const float d1 = z.y - log1pf(fmaxf(target[3 * t + 1], 0.f));
        const float d2 = z.z - log1pf(fmaxf(target[3 * t + 2], 0.f));
        l = d0 * d0 + d1 * d1 + d2 * d2;
        uint16_t* o = g + t * 16;
        const float s = 2.f * inv_count;
        st4(o, s * d0, s * d1, s * d2, 0.f);
        st4(o + 4, 0.f, 0.f, 0.f, 0.f);
        st4(o + 8, 0.f, 0.f, 0.f, 0.f);
        st4(o + 12, 0.f, 0.f, 0.f, 0.f);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, double(red[0]) * double(inv_count));
}

__global__ void k_adam(float* w, float* m, float* v, const float* g, uint16_t* w16, size_t n, float lr, float b1,
                       float b2, float c1, float c2, float eps) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    w16[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* w, uint16_t* w16, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) w16[i] = f2bf(w[i]);
}

// input-gradient weights of one layer: Wt[ci][tap' * cout + co] = W[co][(8 - tap') * cin + ci]
__global__ void k_wflip(const float* w, uint16_t* wt, uint32_t cin, uint32_t cout, uint32_t kpad, uint32_t kpad_t) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(cin) * kpad_t) return;
    const uint32_t ci = uint32_t(t / kpad_t), k = uint32_t(t % kpad_t);
    float v = 0.f;
    if (k < 9u * cout) {
        const uint32_t tap = k / cout, co = k - tap * cout;
        v = w[size_t(co) * kpad + (8u - tap) * cin + ci];
    }
    wt[t] = f2bf(v);
}

// features of the current frame (see gk_denoise_input)
__global__ void k_frame_input(const ushort4* acc_d, const ushort4* acc_s, const ushort4* acc_a, const ushort4* nrm,
                              const float4* phit, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 d = from_h4(acc_d[i]), s = from_h4(acc_s[i]), a = from_h4(acc_a[i]), nm = from_h4(nrm[i]);
    const float t = phit[i].x;
    float* o = out + size_t(i) * kIn;
    o[0] = d.x * a.x + s.x;
    o[1] = d.y * a.y + s.y;
    o[2] = d.z * a.z + s.z;
    o[3] = a.x;
    o[4] = a.y;
    o[5] = a.z;
    o[6] = nm.x;
    o[7] = nm.y;
    o[8] = nm.z;
    o[9] = log1pf(fminf(fmaxf(t, 0.f), 1e4f)) * 0.1f;
}

// ---------------------------------------------------------------- host side
enum Tensor {
    T_X, T_A0, T_S0, T_P0, T_S1, T_P1, T_S2, T_P2, T_S3, T_P3, T_B, T_D3, T_D2, T_D1, T_D0, T_E,
    // gradients
    G_Y, G_E, G_D0, G_U0, G_S0, G_D1, G_U1, G_S1, G_D2, G_U2, G_S2, G_D3, G_U3, G_S3, G_B, G_P3, G_P2, G_P1, G_P0,
    G_A0, T_COUNT
};
constexpr int kNone = -1;

struct LayerSpec {
    int a, b, out;             // activation tensors
    int ga, gb;                // gradient tensors of the inputs (kNone: not needed)
    int gout;                  // gradient tensor of the output
    int post;                  // 0 none, 1 upsum ga -> gup_dst, 2 pool_bwd ga -> pool_dst
    int post_dst;
};

// forward order; pools run after layers 1, 2, 3 and 4
const LayerSpec kSpec[kLayers] = {
    {T_X, kNone, T_A0, kNone, kNone, G_A0, 0, kNone},   // L0
    {T_A0, kNone, T_S0, G_A0, kNone, G_S0, 0, kNone},   // L1
    {T_P0, kNone, T_S1, G_P0, kNone, G_S1, 2, G_S0},    // L2
    {T_P1, kNone, T_S2, G_P1, kNone, G_S2, 2, G_S1},    // L3
    {T_P2, kNone, T_S3, G_P2, kNone, G_S3, 2, G_S2},    // L4
    {T_P3, kNone, T_B, G_P3, kNone, G_B, 2, G_S3},      // L5
    {T_B, T_S3, T_D3, G_U3, G_S3, G_D3, 1, G_B},        // L6
    {T_D3, T_S2, T_D2, G_U2, G_S2, G_D2, 1, G_D3},      // L7
    {T_D2, T_S1, T_D1, G_U1, G_S1, G_D1, 1, G_D2},      // L8
    {T_D1, T_S0, T_D0, G_U0, G_S0, G_D0, 1, G_D1},      // L9
    {T_D0, kNone, T_E, G_D0, kNone, G_E, 0, kNone},     // L10
    {T_E, kNone, kNone, G_E, kNone, G_Y, 0, kNone},     // L11
};

struct TensorInfo {
    int level;
    uint32_t c;
};

inline uint32_t round16(uint32_t c) { return (c + 15u) & ~15u; }
inline uint32_t round32(uint32_t c) { return (c + 31u) & ~31u; }
inline uint32_t nblk(size_t n, uint32_t b = 256) { return uint32_t((n + b - 1) / b); }

// PCG32 (O'Neill) and Box-Muller for the weight initialisation
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed) : state(0), inc((0x853c49e6748fea9bull << 1) | 1u) {
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    double uniform() { return (double(next()) + 0.5) / 4294967296.0; }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace unet
}  // namespace gk

using namespace gk;
using namespace gk::unet;

struct gk_unet {
    Ctx* ctx = nullptr;
    uint32_t base = 0;
    uint32_t ch[5] = {};
    gk_unet_layer L[kLayers] = {};
    uint32_t kpad_t[kLayers] = {};  // row length of the flipped weights
    size_t off_t[kLayers] = {};
    size_t nparam = 0, nparam_t = 0;
    DevBuf<float> w32, m, v, grad;
    DevBuf<uint16_t> w16, wt16;
    DevBuf<double> loss;
    uint64_t step = 0;
    TensorInfo tinfo[T_COUNT] = {};
    DevBuf<uint16_t> t[T_COUNT];
    DevBuf<float> y32;                 // network output, float4 per pixel
    DevBuf<float> feat;                // frame features (denoiser path)
    size_t ws_px = 0;                  // level-0 pixels the workspace holds
    bool ws_grad = false;
    bool fixed_shapes = true;          // forward through k_convs where a shape is built
};

namespace {

void describe(gk_unet& n, uint32_t base) {
    n.base = base;
    const float mul[5] = {1.f, 1.5f, 2.f, 3.f, 4.f};
    for (int l = 0; l < 5; ++l) n.ch[l] = round16(uint32_t(std::ceil(base * mul[l])));
    const uint32_t* c = n.ch;
    // {cin_a, cin_b, cout, up, relu, level}
    const uint32_t d[kLayers][6] = {
        {kInPad, 0, c[0], 0, 1, 0}, {c[0], 0, c[0], 0, 1, 0}, {c[0], 0, c[1], 0, 1, 1}, {c[1], 0, c[2], 0, 1, 2},
        {c[2], 0, c[3], 0, 1, 3},   {c[3], 0, c[4], 0, 1, 4}, {c[4], c[3], c[3], 1, 1, 3}, {c[3], c[2], c[2], 1, 1, 2},
        {c[2], c[1], c[1], 1, 1, 1}, {c[1], c[0], c[0], 1, 1, 0}, {c[0], 0, c[0], 0, 1, 0}, {c[0], 0, 16, 0, 0, 0},
    };
    size_t off = 0, off_t = 0;
    for (uint32_t l = 0; l < kLayers; ++l) {
        gk_unet_layer& L = n.L[l];
        L.cin_a = d[l][0];
        L.cin_b = d[l][1];
        L.cout = d[l][2];
        L.cout_real = l + 1 == kLayers ? 3u : L.cout;
        L.up = d[l][3];
        L.relu = d[l][4];
        L.level = d[l][5];
        const uint32_t cin = L.cin_a + L.cin_b;
        L.kpad = round32(9u * cin + 1u);
        L.offset = off;
        off += size_t(L.cout) * L.kpad;
        n.kpad_t[l] = round32(9u * L.cout);
        n.off_t[l] = off_t;
        off_t += size_t(cin) * n.kpad_t[l];
    }
    n.nparam = off;
    n.nparam_t = off_t;
    const TensorInfo ti[T_COUNT] = {
        {0, kInPad}, {0, c[0]}, {0, c[0]}, {1, c[0]}, {1, c[1]}, {2, c[1]}, {2, c[2]}, {3, c[2]}, {3, c[3]}, {4, c[3]},
        {4, c[4]}, {3, c[3]}, {2, c[2]}, {1, c[1]}, {0, c[0]}, {0, c[0]},
        // gradients
        {0, 16}, {0, c[0]}, {0, c[0]}, {0, c[1]}, {0, c[0]}, {1, c[1]}, {1, c[2]}, {1, c[1]}, {2, c[2]}, {2, c[3]},
        {2, c[2]}, {3, c[3]}, {3, c[4]}, {3, c[3]}, {4, c[4]}, {4, c[3]}, {3, c[2]}, {2, c[1]}, {1, c[0]}, {0, c[0]},
    };
    for (int i = 0; i < T_COUNT; ++i) n.tinfo[i] = ti[i];
}

int refresh_bf16(gk_unet& n) {
    hipStream_t st = n.ctx->stream;
    k_to_bf16<<<nblk(n.nparam), 256, 0, st>>>(n.w32.ptr, n.w16.ptr, n.nparam);
    GK_HIP(hipGetLastError());
    for (uint32_t l = 0; l < kLayers; ++l) {
        const gk_unet_layer& L = n.L[l];
        const uint32_t cin = L.cin_a + L.cin_b;
        k_wflip<<<nblk(size_t(cin) * n.kpad_t[l]), 256, 0, st>>>(n.w32.ptr + L.offset, n.wt16.ptr + n.off_t[l], cin,
                                                                 L.cout, L.kpad, n.kpad_t[l]);
        GK_HIP(hipGetLastError());
    }
    return GK_OK;
}

int ensure_ws(gk_unet& n, uint32_t nimg, uint32_t h, uint32_t w, bool grad) {
    // level l holds (n h w) >> 2l pixels: h and w are multiples of 16 here
    const size_t px0 = size_t(nimg) * h * w;
    if (px0 <= n.ws_px && (!grad || n.ws_grad)) return GK_OK;
    const size_t cap = px0 > n.ws_px ? px0 : n.ws_px;
    const bool with_grad = grad || n.ws_grad;
    for (int i = 0; i < T_COUNT; ++i) {
        if (i >= G_Y && !with_grad) continue;
        GK_HIP(n.t[i].resize((cap >> (2 * n.tinfo[i].level)) * n.tinfo[i].c));
    }
    GK_HIP(n.y32.resize(cap * 4));
    n.ws_px = cap;
    n.ws_grad = with_grad;
    return GK_OK;
}

// LDS a work-group of the current device may allocate (160 KB on gfx950)
static uint32_t max_lds_per_block() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536u;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v < 65536)
        return 65536u;
    return uint32_t(v);
}

// the input tile goes through LDS when it fits a work-group's LDS; above the default 64 KB
// of dynamic LDS the kernel's limit is raised first (the level-3 decoder convolution of the
// 32-channel network stages 86 KB: 76.8 -> 57.9 us at 1080p), and a refused raise keeps the
// direct loads
template <int MT>
void launch_conv_t(const ConvK& K, uint32_t nimg, hipStream_t st) {
    const dim3 grid((K.W + 15) / 16, K.tiles_y * nimg, K.cout / (MT * 16));
    const uint32_t bytes = conv_tile_bytes(K.ca, K.cb, K.a_up);
    const uint32_t wbytes = conv_weight_bytes(MT, K.kpad);
    bool stage = (K.W % 2u) == 0u && (K.H % 2u) == 0u && bytes <= max_lds_per_block();
    // weights join the tile while both stay within 64 KB (two or more work-groups per CU)
    if (stage && bytes + wbytes <= 65536u) {
        k_conv<MT, 2><<<grid, 256, bytes + wbytes, st>>>(K);
        return;
    }
    if (stage && bytes > 65536u) {
        const void* fn = reinterpret_cast<const void*>(&k_conv<MT, 1>);
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)) != hipSuccess) {
            (void)hipGetLastError();
            stage = false;
        }
    }
    if (stage) k_conv<MT, 1><<<grid, 256, bytes, st>>>(K);
    else k_conv<MT, 0><<<grid, 256, 0, st>>>(K);
}

int launch_conv(ConvK K, uint32_t nimg, hipStream_t st) {
    // the most output channels per work-group that still give at least 256 work-groups (the
    // coarse levels have 40 to 135 pixel tiles: one work-group per tile left most CUs idle),
    // unless the staged input tile is above 64 KB (then every extra work-group stages it
    // again: the level-3 decoder convolution's 86 KB tile ran at 72 us split three ways
    // against 58 us whole)
    const uint32_t nm = K.cout / 16;
    K.tiles_y = (K.H + 15) / 16;
    const uint32_t tiles = ((K.W + 15) / 16) * K.tiles_y * nimg;
    const bool split = conv_tile_bytes(K.ca, K.cb, K.a_up) <= 65536u;
    int mt = 1;
    for (int c = 8; c >= 1; --c)
        if (nm % uint32_t(c) == 0u && (!split || tiles * (nm / uint32_t(c)) >= 256u)) {
            mt = c;
            break;
        }
    switch (mt) {
        case 8: launch_conv_t<8>(K, nimg, st); break;
        case 7: launch_conv_t<7>(K, nimg, st); break;
        case 6: launch_conv_t<6>(K, nimg, st); break;
        case 5: launch_conv_t<5>(K, nimg, st); break;
        case 4: launch_conv_t<4>(K, nimg, st); break;
        case 3: launch_conv_t<3>(K, nimg, st); break;
        case 2: launch_conv_t<2>(K, nimg, st); break;
        default: launch_conv_t<1>(K, nimg, st); break;
    }
    GK_HIP(hipGetLastError());
    return GK_OK;
}

// weights in LDS and double buffering while the work-group's LDS stays within 80 KB (two
// work-groups per CU); otherwise weights from L2, or a single buffer
constexpr uint32_t kConvLdsBudget = 80u * 1024u;
template <int CA, int CB, bool UP, int COUT, bool POOL, int OUT, bool WL, int IN, bool DB>
int launch_convs_w(ConvS K, uint32_t nimg, hipStream_t st, bool persistent) {
    using P = ConvP<CA, CB, UP, COUT, POOL, OUT, WL, IN, DB>;
    const uint32_t bytes = uint32_t(P::LDSU) * 16u;
    if (bytes > max_lds_per_block()) return GK_ERR_LIMIT;
    const void* fn = reinterpret_cast<const void*>(&k_convs<CA, CB, UP, COUT, POOL, OUT, WL, IN, DB>);
    if (bytes > 65536u) GK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
    int dev = 0, ncu = 0, per_cu = 0;
    GK_HIP(hipGetDevice(&dev));
    GK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    GK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, bytes));
    K.tiles = uint32_t(nimg) * K.tiles_y * ((K.W + 15) / 16);
    uint32_t G = uint32_t(std::max(1, ncu) * std::max(1, per_cu));
    if (!persistent || G > K.tiles) G = K.tiles;
    k_convs<CA, CB, UP, COUT, POOL, OUT, WL, IN, DB><<<G, 256, bytes, st>>>(K);
    GK_HIP(hipGetLastError());
    return GK_OK;
}
// mode 0: persistent, weights in LDS and double buffering (80 KB at most); mode 1: one tile
// per work-group, weights read from L2; mode 2: one tile per work-group, weights in LDS;
// mode 3: persistent, weights in LDS, one buffer.
// Chosen per layer by measurement at 1920x1088 (the level-0 decoder convolution's 47 KB of
// weights leave no room for two buffers and two work-groups per CU; the computed-input first
// layer and the level-2 encoder convolution gain nothing from persistence)
template <int CA, int CB, bool UP, int COUT, bool POOL, bool OUTF, bool WL, int IN, int R = 4>
int launch_convs1(ConvS K, uint32_t nimg, hipStream_t st) {
    using S = ConvShape<CA, CB, UP, R>;
    const uint32_t in_bytes = uint32_t(S::UNITS) * 16u + (WL ? uint32_t(COUT) * uint32_t(S::KPAD / 8 + 1) * 16u : 0u);
    const uint32_t out_bytes = 64u * R * (OUTF ? 16u : uint32_t(COUT) * 2u);
    const uint32_t bytes = in_bytes > out_bytes ? in_bytes : out_bytes;
    if (bytes > max_lds_per_block()) return GK_ERR_LIMIT;
    const void* fn = reinterpret_cast<const void*>(&k_convs1<CA, CB, UP, COUT, POOL, OUTF, WL, IN, R>);
    if (bytes > 65536u) GK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
    K.tiles_y = (K.H + 4u * R - 1u) / (4u * R);
    const dim3 grid((K.W + 15) / 16, K.tiles_y * nimg, 1);
    k_convs1<CA, CB, UP, COUT, POOL, OUTF, WL, IN, R><<<grid, 256, bytes, st>>>(K);
    GK_HIP(hipGetLastError());
    return GK_OK;
}
template <int CA, int CB, bool UP, int COUT, bool POOL, int OUT, int IN = 0>
int launch_convs_t(const ConvS& K, uint32_t nimg, hipStream_t st, int mode = 0) {
    static_assert(OUT <= 1 || IN == 0, "");
    if (mode == 1) return launch_convs1<CA, CB, UP, COUT, POOL, (OUT != 0), false, IN>(K, nimg, st);
    if (mode == 2) return launch_convs1<CA, CB, UP, COUT, POOL, (OUT != 0), true, IN>(K, nimg, st);
    if (mode == 3) return launch_convs_w<CA, CB, UP, COUT, POOL, OUT, true, IN, false>(K, nimg, st, true);
    if (ConvP<CA, CB, UP, COUT, POOL, OUT, true, IN, true>::LDSU * 16u <= kConvLdsBudget)
        return launch_convs_w<CA, CB, UP, COUT, POOL, OUT, true, IN, true>(K, nimg, st, true);
    return launch_convs_w<CA, CB, UP, COUT, POOL, OUT, false, IN, true>(K, nimg, st, true);
}

// the forward convolution of layer l through its fixed-shape kernel when one is built for
// its channel counts (the 32-channel network), with the following 2 x 2 pool fused when
// pool is set; kNoShape otherwise (the caller falls back to k_conv)
constexpr int kNoShape = 1;
int launch_convs(const gk_unet_layer& L, ConvS K, uint32_t nimg, bool pool, int in_mode, int out_mode,
                 hipStream_t st) {
    K.tiles_y = (K.H + 15) / 16;
    if (in_mode != 0 || out_mode >= 2) {
        if (nimg != 1) return kNoShape;
        if (in_mode != 0) {
            if (!(L.cin_a == 16 && L.cin_b == 0 && !L.up && L.cout == 32 && !pool)) return kNoShape;
            return in_mode == 1 ? launch_convs_t<16, 0, false, 32, false, 0, 1>(K, nimg, st, 2)
                                : launch_convs_t<16, 0, false, 32, false, 0, 2>(K, nimg, st, 2);
        }
        if (!(L.cin_a == 32 && L.cin_b == 0 && !L.up && L.cout == 16 && !pool)) return kNoShape;
        return out_mode == 2 ? launch_convs_t<32, 0, false, 16, false, 2>(K, nimg, st)
                             : launch_convs_t<32, 0, false, 16, false, 3>(K, nimg, st);
    }
    const bool outf = K.outf != nullptr;
    const uint32_t key[6] = {L.cin_a, L.cin_b, L.up, L.cout, pool ? 1u : 0u, outf ? 1u : 0u};
    auto is = [&](uint32_t a, uint32_t b, uint32_t up, uint32_t co, uint32_t pl, uint32_t of) {
        return key[0] == a && key[1] == b && key[2] == up && key[3] == co && key[4] == pl && key[5] == of;
    };
    // built where they beat k_conv (+ k_pool) at 1920x1088: the coarse levels' wide layers
    // (level 3-4, and the decoder's 112- to 224-channel inputs) run faster through k_conv's
    // rolled K loop, whose registers leave room for more waves
    if (is(16, 0, 0, 32, 0, 0)) return launch_convs_t<16, 0, false, 32, false, 0>(K, nimg, st, 2);
    if (is(32, 0, 0, 32, 1, 0)) return launch_convs_t<32, 0, false, 32, true, 0>(K, nimg, st);
    if (is(32, 0, 0, 48, 1, 0)) return launch_convs_t<32, 0, false, 48, true, 0>(K, nimg, st);
    if (is(48, 0, 0, 64, 1, 0)) return launch_convs_t<48, 0, false, 64, true, 0>(K, nimg, st, 1);
    if (is(48, 32, 1, 32, 0, 0)) return launch_convs_t<48, 32, true, 32, false, 0>(K, nimg, st, 1);
    if (is(64, 48, 1, 48, 0, 0)) return launch_convs_t<64, 48, true, 48, false, 0>(K, nimg, st, 1);
    if (is(32, 0, 0, 32, 0, 0)) return launch_convs_t<32, 0, false, 32, false, 0>(K, nimg, st);
    if (is(32, 0, 0, 16, 0, 1)) return launch_convs_t<32, 0, false, 16, false, 1>(K, nimg, st);
    return kNoShape;
}

template <int MT, int NT>
void launch_wgrad_t(const WgradK& K, hipStream_t st) {
    const dim3 grid((K.rows + K.rows_per_block - 1) / K.rows_per_block, K.kpad / (NT * 16), K.cout / (MT * 16));
    k_wgrad<MT, NT><<<grid, 256, 0, st>>>(K);
}

int launch_wgrad(WgradK K, hipStream_t st) {
    const uint32_t nm = K.cout / 16, nn = K.kpad / 16;
    const int mt = nm % 4u == 0u ? 4 : (nm % 3u == 0u ? 3 : (nm % 2u == 0u ? 2 : 1));
    const int nt = nn % 4u == 0u ? 4 : 2;
    K.rows_per_block = 16;
    switch (mt * 10 + nt) {
        case 44: launch_wgrad_t<4, 4>(K, st); break;
        case 42: launch_wgrad_t<4, 2>(K, st); break;
        case 34: launch_wgrad_t<3, 4>(K, st); break;
        case 32: launch_wgrad_t<3, 2>(K, st); break;
        case 24: launch_wgrad_t<2, 4>(K, st); break;
        case 22: launch_wgrad_t<2, 2>(K, st); break;
        case 14: launch_wgrad_t<1, 4>(K, st); break;
        default: launch_wgrad_t<1, 2>(K, st); break;
    }
    GK_HIP(hipGetLastError());
    return GK_OK;
}

const uint16_t* tp(gk_unet& n, int i) { return i == kNone ? nullptr : n.t[i].ptr; }

// forward pass over X (already packed) at n x h x w; output in y32
// where the network's input comes from and its output goes: by default the packed tensor
// T_X and y32; the inference paths instead compute the input inside the first convolution
// (from FP32 features or the frame's targets) and write the radiance or the display image
// from the last one (fused_io)
struct FwdIO {
    int in_mode = 0;                  // 0 T_X, 1 feat, 2 frame targets
    int out_mode = 1;                 // 1 y32, 2 orgb, 3 odisp
    const float* feat = nullptr;
    const ushort4 *fd = nullptr, *fs = nullptr, *fa = nullptr, *fn = nullptr;
    const float4* fh = nullptr;
    float* orgb = nullptr;
    ushort4* odisp = nullptr;
    uint32_t vw = 0, vh = 0;
    float paper_white = 0.f;
    uint32_t hdr = 0, linear = 0;
};

// the fused first and last layers exist for this network (32 channels, one image)
bool fused_io(const gk_unet& n, uint32_t nimg) {
    const gk_unet_layer &a = n.L[0], &z = n.L[kLayers - 1];
    return n.fixed_shapes && nimg == 1 && a.cin_a == 16 && a.cin_b == 0 && a.cout == 32 && z.cin_a == 32 &&
           z.cin_b == 0 && z.cout == 16;
}

int forward(gk_unet& n, uint32_t nimg, uint32_t h, uint32_t w, const FwdIO& io = FwdIO{}) {
    hipStream_t st = n.ctx->stream;
    for (uint32_t l = 0; l < kLayers; ++l) {
        const gk_unet_layer& L = n.L[l];
        const LayerSpec& S = kSpec[l];
        const bool pool = l >= 1 && l <= 4;  // pool the skip output to the next level
        const uint32_t lw = w >> L.level, lh = h >> L.level;
        const int in_mode = l == 0 ? io.in_mode : 0;
        const int out_mode = l + 1 == kLayers ? io.out_mode : 0;
        if (in_mode != 0 || out_mode >= 2) {  // fused_io() checked by the caller
            ConvS Q{};
            Q.a = tp(n, S.a);
            Q.w = n.w16.ptr + L.offset;
            Q.W = lw;
            Q.H = lh;
            Q.relu = L.relu;
            Q.out = S.out == kNone ? nullptr : n.t[S.out].ptr;
            Q.feat = io.feat;
            Q.fd = io.fd;
            Q.fs = io.fs;
            Q.fa = io.fa;
            Q.fn = io.fn;
            Q.fh = io.fh;
            Q.orgb = io.orgb;
            Q.odisp = io.odisp;
            Q.vw = io.vw;
            Q.vh = io.vh;
            Q.paper_white = io.paper_white;
            Q.hdr = io.hdr;
            Q.linear = io.linear;
            const int s = launch_convs(L, Q, nimg, pool, in_mode, out_mode, st);
            if (s != GK_OK) return s == kNoShape ? GK_ERR_STATE : s;
            continue;
        }
        if (n.fixed_shapes && (!(pool || L.up) || ((lw % 2u) == 0u && (lh % 2u) == 0u))) {
            ConvS Q{};
            Q.a = tp(n, S.a);
            Q.b = tp(n, S.b);
            Q.w = n.w16.ptr + L.offset;
            Q.W = lw;
            Q.H = lh;
            Q.relu = L.relu;
            if (S.out == kNone) Q.outf = n.y32.ptr;
            else Q.out = n.t[S.out].ptr;
            if (pool) Q.pool = n.t[kSpec[l + 1].a].ptr;
            const int s = launch_convs(L, Q, nimg, pool, 0, 0, st);
            if (s == GK_OK) continue;
            if (s != kNoShape) return s;
        }
        ConvK K{};
        K.a = tp(n, S.a);
        K.b = tp(n, S.b);
        K.w = n.w16.ptr + L.offset;
        K.ca = L.cin_a;
        K.cb = L.cin_b;
        K.cin = L.cin_a + L.cin_b;
        K.a_up = L.up;
        K.W = w >> L.level;
        K.H = h >> L.level;
        K.kpad = L.kpad;
        K.ktap = 9u * K.cin;
        K.cout = L.cout;
        K.split = L.cout;
        K.relu = L.relu;
        K.bias = 1;
        if (S.out == kNone) K.outf = n.y32.ptr;
        else K.out = n.t[S.out].ptr;
        int s = launch_conv(K, nimg, st);
        if (s != GK_OK) return s;
        if (pool) {
            const int src = S.out, dst = kSpec[l + 1].a;
            const size_t cnt = size_t(nimg) * (lh / 2) * (lw / 2) * (n.tinfo[src].c / 8);
            k_pool<<<nblk(cnt), 256, 0, st>>>(n.t[src].ptr, n.t[dst].ptr, nimg, lh, lw, n.tinfo[src].c);
            GK_HIP(hipGetLastError());
        }
    }
    return GK_OK;
}

int pack_input(gk_unet& n, const float* in, uint32_t nimg, uint32_t h, uint32_t w, uint32_t hp, uint32_t wp) {
    const size_t cnt = size_t(nimg) * hp * wp;
    k_pack<<<nblk(cnt), 256, 0, n.ctx->stream>>>(in, n.t[T_X].ptr, nimg, h, w, hp, wp);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

int backward(gk_unet& n, const float* target, uint32_t nimg, uint32_t h, uint32_t w, float* grad) {
    hipStream_t st = n.ctx->stream;
    const size_t npx = size_t(nimg) * h * w;
    k_fill<<<nblk(n.nparam), 256, 0, st>>>(grad, n.nparam);
    GK_HIP(hipGetLastError());
    k_fill<<<1, 32, 0, st>>>(reinterpret_cast<float*>(n.loss.ptr), 2);
    GK_HIP(hipGetLastError());
    k_loss<<<nblk(npx), 256, 0, st>>>(n.y32.ptr, target, n.t[G_Y].ptr, npx, 1.f / float(3 * npx), n.loss.ptr);
    GK_HIP(hipGetLastError());
    for (int l = int(kLayers) - 1; l >= 0; --l) {
        const gk_unet_layer& L = n.L[l];
        const LayerSpec& S = kSpec[l];
        const uint32_t lw = w >> L.level, lh = h >> L.level;
        const uint32_t cin = L.cin_a + L.cin_b;
        uint16_t* g = n.t[S.gout].ptr;
        if (L.relu) {
            const size_t n8 = size_t(nimg) * lh * lw * L.cout / 8;
            k_relu_mask<<<nblk(n8), 256, 0, st>>>(g, n.t[S.out].ptr, n8);
            GK_HIP(hipGetLastError());
        }
        WgradK Wk{};
        Wk.a = tp(n, S.a);
        Wk.b = tp(n, S.b);
        Wk.g = g;
        Wk.dw = grad + L.offset;
        Wk.ca = L.cin_a;
        Wk.cb = L.cin_b;
        Wk.cin = cin;
        Wk.a_up = L.up;
        Wk.W = lw;
        Wk.H = lh;
        Wk.rows = nimg * lh;
        Wk.kpad = L.kpad;
        Wk.ktap = 9u * cin;
        Wk.cout = L.cout;
        Wk.bias = 1;
        int s = launch_wgrad(Wk, st);
        if (s != GK_OK) return s;
        if (S.ga == kNone) continue;
        ConvK K{};
        K.a = g;
        K.ca = L.cout;
        K.cb = 0;
        K.cin = L.cout;
        K.a_up = 0;
        K.w = n.wt16.ptr + n.off_t[l];
        K.W = lw;
        K.H = lh;
        K.kpad = n.kpad_t[l];
        K.ktap = 9u * L.cout;
        K.cout = cin;
        K.split = L.cin_a;
        K.relu = 0;
        K.bias = 0;
        K.out = n.t[S.ga].ptr;
        K.out2 = S.gb == kNone ? nullptr : n.t[S.gb].ptr;
        s = launch_conv(K, nimg, st);
        if (s != GK_OK) return s;
        if (S.post == 1) {  // gradient of the upsampled input -> low resolution
            const uint32_t c = L.cin_a;
            const size_t cnt = size_t(nimg) * (lh / 2) * (lw / 2) * (c / 8);
            k_upsum<<<nblk(cnt), 256, 0, st>>>(n.t[S.ga].ptr, n.t[S.post_dst].ptr, nimg, lh, lw, c);
            GK_HIP(hipGetLastError());
        } else if (S.post == 2) {  // gradient of the pooled input -> the skip tensor above
            const uint32_t c = L.cin_a;
            const uint32_t uh = lh * 2, uw = lw * 2;
            const size_t cnt = size_t(nimg) * lh * lw * (c / 8);
            k_pool_bwd<<<nblk(cnt), 256, 0, st>>>(n.t[S.ga].ptr, n.t[S.post_dst].ptr, nimg, uh, uw, c);
            GK_HIP(hipGetLastError());
        }
    }
    return GK_OK;
}

int check_train_args(gk_unet* net, const float* in, const float* tg, uint32_t nimg, uint32_t h, uint32_t w) {
    if (!net || !net->ctx || !in || !tg || nimg == 0 || h == 0 || w == 0) return GK_ERR_ARG;
    if (h % 16u || w % 16u) return GK_ERR_ARG;
    if (size_t(nimg) * h * w > (size_t(1) << 27)) return GK_ERR_LIMIT;
    return GK_OK;
}

int grad_impl(gk_unet& n, const float* in, const float* tg, uint32_t nimg, uint32_t h, uint32_t w, float* grad,
              float* loss) {
    int s = ensure_ws(n, nimg, h, w, true);
    if (s != GK_OK) return s;
    if ((s = pack_input(n, in, nimg, h, w, h, w)) != GK_OK) return s;
    if ((s = forward(n, nimg, h, w)) != GK_OK) return s;
    if ((s = backward(n, tg, nimg, h, w, grad)) != GK_OK) return s;
    if (loss) {
        double l = 0.0;
        GK_HIP(hipMemcpyAsync(&l, n.loss.ptr, sizeof(double), hipMemcpyDeviceToHost, n.ctx->stream));
        GK_HIP(hipStreamSynchronize(n.ctx->stream));
        *loss = float(l);
    }
    return GK_OK;
}

}  // namespace

namespace gk {

// GK_DENOISER_UNET: features of the frame -> network -> tonemapped GK_RT_DENOISED
int unet_denoise_frame(Ctx& c) {
    if (!c.unet) return GK_ERR_STATE;
    gk_unet& n = *c.unet;
    const uint32_t W = c.width, H = c.height, n_px = W * H;
    const uint32_t wp = (W + 15u) & ~15u, hp = (H + 15u) & ~15u;
    int s = ensure_ws(n, 1, hp, wp, false);
    if (s != GK_OK) return s;
    if (fused_io(n, 1)) {  // features computed in the first convolution, display from the last
        FwdIO io;
        io.in_mode = 2;
        io.out_mode = 3;
        io.fd = static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_DIFFUSE]);
        io.fs = static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_SPECULAR]);
        io.fa = static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_ALBEDO]);
        io.fn = static_cast<const ushort4*>(c.target_ptr[GK_RT_NORMAL]);
        io.fh = static_cast<const float4*>(c.target_ptr[GK_RT_PRIMARY_HIT]);
        io.odisp = static_cast<ushort4*>(c.target_ptr[GK_RT_DENOISED]);
        io.vw = W;
        io.vh = H;
        io.paper_white = c.params.paper_white_nit;
        io.hdr = c.params.hdr ? 1u : 0u;
        io.linear = (c.params.flags & GK_FRAME_LINEAR_OUTPUT) ? 1u : 0u;
        return forward(n, 1, hp, wp, io);
    }
    if (n.feat.count < size_t(n_px) * kIn) GK_HIP(n.feat.resize(size_t(n_px) * kIn));
    k_frame_input<<<nblk(n_px), 256, 0, c.stream>>>(
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_DIFFUSE]),
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_SPECULAR]),
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_ALBEDO]), static_cast<const ushort4*>(c.target_ptr[GK_RT_NORMAL]),
        static_cast<const float4*>(c.target_ptr[GK_RT_PRIMARY_HIT]), n.feat.ptr, n_px);
    GK_HIP(hipGetLastError());
    if ((s = pack_input(n, n.feat.ptr, 1, H, W, hp, wp)) != GK_OK) return s;
    if ((s = forward(n, 1, hp, wp)) != GK_OK) return s;
    k_unpack_display<<<nblk(n_px), 256, 0, c.stream>>>(n.y32.ptr, static_cast<ushort4*>(c.target_ptr[GK_RT_DENOISED]),
                                                       H, W, wp, c.params.paper_white_nit, c.params.hdr ? 1u : 0u,
                                                       (c.params.flags & GK_FRAME_LINEAR_OUTPUT) ? 1u : 0u);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

}  // namespace gk

extern "C" {

int gk_unet_create(gk_ctx* ctx, uint32_t base_channels, uint64_t seed, gk_unet** out) {
    if (!ctx || !out || base_channels < 8u || base_channels > 256u) return GK_ERR_ARG;
    ctx->c.early_ok = false;  // queues work on the context's stream
    *out = nullptr;
    Ctx& c = ctx->c;
    GK_HIP(hipSetDevice(c.device));
    gk_unet* n = new gk_unet();
    n->ctx = &c;
    describe(*n, base_channels);
    const char* generic = std::getenv("GKNX_UNET_GENERIC");
    n->fixed_shapes = !(generic && generic[0] && generic[0] != '0');
    auto fail = [&](int s) {
        n->ctx = nullptr;
        delete n;
        return s;
    };
    if (n->w32.resize(n->nparam) != hipSuccess || n->m.resize(n->nparam) != hipSuccess ||
        n->v.resize(n->nparam) != hipSuccess || n->grad.resize(n->nparam) != hipSuccess ||
        n->w16.resize(n->nparam) != hipSuccess || n->wt16.resize(n->nparam_t) != hipSuccess ||
        n->loss.resize(1) != hipSuccess)
        return fail(GK_ERR_OOM);
    // He-normal weights over the real inputs of each output channel, zero bias and padding
    std::vector<float> h(n->nparam, 0.f);
    Pcg32 rng(seed);
    for (uint32_t l = 0; l < kLayers; ++l) {
        const gk_unet_layer& L = n->L[l];
        const uint32_t cin = L.cin_a + L.cin_b;
        const uint32_t cin_real = l == 0 ? kIn : cin;
        const double sd = std::sqrt((l + 1 == kLayers ? 1.0 : 2.0) / (9.0 * cin_real));
        for (uint32_t co = 0; co < L.cout_real; ++co)
            for (uint32_t tap = 0; tap < 9; ++tap)
                for (uint32_t ci = 0; ci < cin_real; ++ci)
                    h[L.offset + size_t(co) * L.kpad + tap * cin + ci] = float(sd * rng.normal());
    }
    hipStream_t st = c.stream;
    if (hipMemcpyAsync(n->w32.ptr, h.data(), n->nparam * 4, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(GK_ERR_HIP);
    k_fill<<<nblk(n->nparam), 256, 0, st>>>(n->m.ptr, n->nparam);
    k_fill<<<nblk(n->nparam), 256, 0, st>>>(n->v.ptr, n->nparam);
    if (hipGetLastError() != hipSuccess) return fail(GK_ERR_HIP);
    if (refresh_bf16(*n) != GK_OK) return fail(GK_ERR_HIP);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(GK_ERR_HIP);
    if (c.unet) (void)gk_unet_destroy(c.unet);
    c.unet = n;
    *out = n;
    return GK_OK;
}

int gk_unet_destroy(gk_unet* net) {
    if (!net) return GK_ERR_ARG;
    if (net->ctx) {
        (void)hipSetDevice(net->ctx->device);
        (void)hipStreamSynchronize(net->ctx->stream);
        if (net->ctx->unet == net) net->ctx->unet = nullptr;
    }
    delete net;
    return GK_OK;
}

int gk_unet_num_params(gk_unet* net, uint64_t* out) {
    if (!net || !out) return GK_ERR_ARG;
    *out = net->nparam;
    return GK_OK;
}

int gk_unet_layer_info(gk_unet* net, uint32_t layer, gk_unet_layer* out) {
    if (!net || !out || layer >= kLayers) return GK_ERR_ARG;
    *out = net->L[layer];
    return GK_OK;
}

int gk_unet_params(gk_unet* net, float** d_params) {
    if (!net || !d_params) return GK_ERR_ARG;
    *d_params = net->w32.ptr;
    return GK_OK;
}

int gk_unet_params_updated(gk_unet* net) {
    if (!net || !net->ctx) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(net->ctx->device));
    return refresh_bf16(*net);
}

int gk_unet_save(gk_unet* net, const char* path) {
    if (!net || !net->ctx || !path) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(net->ctx->device));
    std::vector<float> h(net->nparam);
    GK_HIP(hipMemcpyAsync(h.data(), net->w32.ptr, net->nparam * 4, hipMemcpyDeviceToHost, net->ctx->stream));
    GK_HIP(hipStreamSynchronize(net->ctx->stream));
    FILE* f = std::fopen(path, "wb");
    if (!f) return GK_ERR_IO;
    const char magic[8] = {'G', 'K', 'U', 'N', 'E', 'T', '0', '1'};
    const uint32_t hdr[2] = {net->base, 0u};
    const uint64_t cnt[2] = {uint64_t(net->nparam), net->step};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(cnt, 8, 2, f) == 2 &&
              std::fwrite(h.data(), 4, h.size(), f) == h.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? GK_OK : GK_ERR_IO;
}

int gk_unet_load(gk_unet* net, const char* path) {
    if (!net || !net->ctx || !path) return GK_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return GK_ERR_IO;
    char magic[8];
    uint32_t hdr[2];
    uint64_t cnt[2];
    std::vector<float> h;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "GKUNET01", 8) == 0 &&
              std::fread(hdr, 4, 2, f) == 2 && std::fread(cnt, 8, 2, f) == 2 && hdr[0] == net->base &&
              cnt[0] == net->nparam;
    if (ok) {
        h.resize(net->nparam);
        ok = std::fread(h.data(), 4, h.size(), f) == h.size();
    }
    std::fclose(f);
    if (!ok) return GK_ERR_IO;
    GK_HIP(hipSetDevice(net->ctx->device));
    hipStream_t st = net->ctx->stream;
    GK_HIP(hipMemcpyAsync(net->w32.ptr, h.data(), net->nparam * 4, hipMemcpyHostToDevice, st));
    k_fill<<<nblk(net->nparam), 256, 0, st>>>(net->m.ptr, net->nparam);
    k_fill<<<nblk(net->nparam), 256, 0, st>>>(net->v.ptr, net->nparam);
    GK_HIP(hipGetLastError());
    net->step = cnt[1];
    const int s = refresh_bf16(*net);
    if (s != GK_OK) return s;
    GK_HIP(hipStreamSynchronize(st));
    return GK_OK;
}

int gk_unet_grad(gk_unet* net, const float* d_input, const float* d_target, uint32_t n, uint32_t h, uint32_t w,
                 float* d_grad, float* loss) {
    int s = check_train_args(net, d_input, d_target, n, h, w);
    if (s != GK_OK) return s;
    if (!d_grad) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(net->ctx->device));
    return grad_impl(*net, d_input, d_target, n, h, w, d_grad, loss);
}

int gk_unet_train_step(gk_unet* net, const float* d_input, const float* d_target, uint32_t n, uint32_t h, uint32_t w,
                       float lr, float* loss) {
    int s = check_train_args(net, d_input, d_target, n, h, w);
    if (s != GK_OK) return s;
    if (!(lr > 0.f)) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(net->ctx->device));
    if ((s = grad_impl(*net, d_input, d_target, n, h, w, net->grad.ptr, loss)) != GK_OK) return s;
    net->step += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = float(1.0 / (1.0 - std::pow(b1, double(net->step))));
    const float c2 = float(1.0 / (1.0 - std::pow(b2, double(net->step))));
    hipStream_t st = net->ctx->stream;
    k_adam<<<nblk(net->nparam), 256, 0, st>>>(net->w32.ptr, net->m.ptr, net->v.ptr, net->grad.ptr, net->w16.ptr,
                                               net->nparam, lr, float(b1), float(b2), c1, c2, 1e-8f);
    GK_HIP(hipGetLastError());
    for (uint32_t l = 0; l < kLayers; ++l) {
        const gk_unet_layer& L = net->L[l];
        const uint32_t cin = L.cin_a + L.cin_b;
        k_wflip<<<nblk(size_t(cin) * net->kpad_t[l]), 256, 0, st>>>(net->w32.ptr + L.offset, net->wt16.ptr + net->off_t[l],
                                                                    cin, L.cout, L.kpad, net->kpad_t[l]);
        GK_HIP(hipGetLastError());
    }
    return GK_OK;
}

int gk_unet_infer(gk_unet* net, const float* d_input, float* d_output, uint32_t h, uint32_t w) {
    if (!net || !net->ctx || !d_input || !d_output || h == 0 || w == 0) return GK_ERR_ARG;
    if (size_t(h) * w > (size_t(1) << 26)) return GK_ERR_LIMIT;
    GK_HIP(hipSetDevice(net->ctx->device));
    const uint32_t hp = (h + 15u) & ~15u, wp = (w + 15u) & ~15u;
    int s = ensure_ws(*net, 1, hp, wp, false);
    if (s != GK_OK) return s;
    if (fused_io(*net, 1)) {  // input computed in the first convolution, radiance from the last
        FwdIO io;
        io.in_mode = 1;
        io.out_mode = 2;
        io.feat = d_input;
        io.orgb = d_output;
        io.vw = w;
        io.vh = h;
        return forward(*net, 1, hp, wp, io);
    }
    if ((s = pack_input(*net, d_input, 1, h, w, hp, wp)) != GK_OK) return s;
    if ((s = forward(*net, 1, hp, wp)) != GK_OK) return s;
    k_unpack<<<nblk(size_t(h) * w), 256, 0, net->ctx->stream>>>(net->y32.ptr, d_output, 1, h, w, hp, wp);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

int gk_denoise_input(gk_ctx* ctx, float* d_input) {
    if (!ctx || !d_input) return GK_ERR_ARG;
    ctx->c.early_ok = false;  // queues work on the context's stream
    Ctx& c = ctx->c;
    if (!c.frame_ready) return GK_ERR_STATE;
    GK_HIP(hipSetDevice(c.device));
    const uint32_t n_px = c.width * c.height;
    k_frame_input<<<nblk(n_px), 256, 0, c.stream>>>(
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_DIFFUSE]),
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_SPECULAR]),
        static_cast<const ushort4*>(c.target_ptr[GK_RT_ACCUM_ALBEDO]), static_cast<const ushort4*>(c.target_ptr[GK_RT_NORMAL]),
        static_cast<const float4*>(c.target_ptr[GK_RT_PRIMARY_HIT]), d_input, n_px);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

}  // extern "C"
