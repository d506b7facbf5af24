// Learned preconditioner: a 3-level 3D convolutional U-Net (channels 8/16/32,
// 3x3x3 kernels) applied to a PCG residual, with its training kernels.
//
// Architecture (fixed; about 66k parameters):
//   L1 (fine):  x0 = [r/rms, fluid, air, solid, 0...]   8 ch   (features)
//               h1 = relu(c1(x0))                         8 ch
//   L2:         h2 = relu(c2(pool(h1)))                  16 ch
//   L3:         h3a = relu(c3(pool(h2)))                 32 ch
//               h3 = relu(c4(h3a))                       32 ch
//   L2:         s2 = relu(c5(up(h3))) + h2               16 ch
//   L1:         s1 = relu(c6(up(s2))) + h1                8 ch
//               d1 = relu(c7(s1))                         8 ch
//               y  = c8(d1)                               1 ch (FP32)
//   z = rms(r) * y on fluid cells, 0 elsewhere.
// pool = 2x2x2 average over the children that exist (ceil sizes), up =
// nearest-neighbour.  Without the biases the map is positively homogeneous,
// and r is normalised by its RMS so the network always sees O(1) inputs.
//
// Data layout: activations are channels-last BF16, cell (i,j,k) of a level at
// ((i*ny + j)*nz + k)*C, so one tap of one cell is C contiguous values
// (16-64 bytes).  Convolutions are implicit GEMMs on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, in the transposed form
// Y^T[Cout x cells] = W^T[Cout x 27*Cin] * X^T[27*Cin x cells]: the A operand
// is the weight matrix, pre-packed once per weight update into the exact
// per-lane fragment order (16 contiguous bytes per lane), and the B operand is
// read from a workgroup's input block staged in LDS — lane (16*g + c) takes
// the 8 channels of tap/channel-group g for cell c.  A wave computes 16
// output channels for 4 rows x 16 consecutive k cells per step.  The input-gradient
// convolution is the same kernel with flipped, transposed weights; weight
// gradients are FP32 outer-product reductions with deterministic partials.
#include "boilsim_hip.h"
#include "grid.hpp"

#include <math.h>
#include <string.h>

namespace bs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Network description
// ---------------------------------------------------------------------------
struct ConvDef {
    int cin, cout, level;
};
constexpr int kNumConv = 8;
__host__ __device__ constexpr ConvDef conv_def(int l) {
    return l == 0 ? ConvDef{8, 8, 0}
         : l == 1 ? ConvDef{8, 16, 1}
         : l == 2 ? ConvDef{16, 32, 2}
         : l == 3 ? ConvDef{32, 32, 2}
         : l == 4 ? ConvDef{32, 16, 1}
         : l == 5 ? ConvDef{16, 8, 0}
         : l == 6 ? ConvDef{8, 8, 0}
                  : ConvDef{8, 1, 0};
}
// FP32 parameter layout: for each conv, W[27][cin][cout] then bias[cout].
__host__ __device__ constexpr size_t conv_w_off(int l) {
    size_t o = 0;
    for (int q = 0; q < l; ++q) o += (size_t)27 * conv_def(q).cin * conv_def(q).cout + conv_def(q).cout;
    return o;
}
__host__ __device__ constexpr size_t conv_b_off(int l) {
    return conv_w_off(l) + (size_t)27 * conv_def(l).cin * conv_def(l).cout;
}
constexpr size_t kNumParams = conv_w_off(kNumConv);

// Packed BF16 fragments: forward and (flipped, transposed) backward per conv.
// A fragment block holds ksteps * mtiles * 64 lanes * 8 values.
__host__ __device__ constexpr int ksteps_of(int cin) { return (27 * cin + 31) / 32; }
__host__ __device__ constexpr int mtiles_of(int cout) { return (cout + 15) / 16; }
__host__ __device__ constexpr size_t frag_elems(int cin, int cout) {
    return (size_t)ksteps_of(cin) * mtiles_of(cout) * 64 * 8;
}
__host__ __device__ constexpr size_t frag_off(int l, int bwd) {
    size_t o = 0;
    for (int q = 0; q < kNumConv; ++q) {
        for (int b = 0; b < 2; ++b) {
            if (q == l && b == bwd) return o;
            const ConvDef d = conv_def(q);
            o += b == 0 ? frag_elems(d.cin, d.cout) : frag_elems(d.cout < 8 ? 8 : d.cout, d.cin);
        }
    }
    return o;
}
constexpr size_t kPackedElems = frag_off(kNumConv, 0);

// ---------------------------------------------------------------------------
// Level geometry
// ---------------------------------------------------------------------------
struct Lvl {
    int nx, ny, nz;
    int64_t n;
    FastDiv fz, fyz;  // division by nz and by ny*nz
};

inline Lvl make_lvl(int nx, int ny, int nz) {
    Lvl L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.n = (int64_t)nx * ny * nz;
    L.fz = make_fastdiv((uint32_t)nz);
    L.fyz = make_fastdiv((uint32_t)(ny * nz));
    return L;
}

__device__ __forceinline__ void cell_ijk(const Lvl& L, uint32_t c, int& i, int& j, int& k) {
    i = (int)fdiv(c, L.fyz);
    const uint32_t rem = c - (uint32_t)i * (uint32_t)(L.ny * L.nz);
    j = (int)fdiv(rem, L.fz);
    k = (int)(rem - (uint32_t)j * (uint32_t)L.nz);
}

// ---------------------------------------------------------------------------
// Implicit-GEMM convolution on MFMA
//   out[cell][co] = act(sum_{t,ci} in[nbr(cell,t)][ci] * W[t][ci][co] + b[co])
//   (+ skip[cell][co] into out_sum when given)
// in is zero outside the grid.
//
// A workgroup owns a 4 x 4 x 16 block of cells (i, j, k): wave w the plane
// i0 + w, its four 16-cell MFMA tiles the rows j0..j0+3.  The block's input
// with a one-cell halo (6 x 6 x 18 cells, zeros outside the grid) is staged
// in LDS once, so the 27 taps of every cell are read from LDS instead of
// being gathered from L1/L2 27 times (the gathers bounded the earlier
// version).  LDS layout [channel group][halo cell][8 channels]: the 16 lanes
// of one B fragment read 16 consecutive 16-byte slots, without bank
// conflicts for any channel count.
// ---------------------------------------------------------------------------
constexpr int kTI = 4, kTJ = 4, kTK = 16;
constexpr int kHI = kTI + 2, kHJ = kTJ + 2, kHK = kTK + 2;
constexpr int kHalo = kHI * kHJ * kHK;

// LDS slot (within one channel group's halo block) of tap t = 9 (di + 1) +
// 3 (dj + 1) + (dk + 1) of the cell at (w, 0, c) of the block, relative to
// that cell's own slot w kHJ kHK + c: halo cell (w + 1 + di, 1 + dj,
// c + 1 + dk).  t >= 27 (the K padding) maps to tap 26, which keeps the
// discarded read inside the tile.
__host__ __device__ constexpr int tap_lds(int t) {
    return t >= 27 ? tap_lds(26) : ((t / 9) * kHJ + (t / 3) % 3) * kHK + t % 3;
}

// tap_lds of taps t0 .. t0 + n - 1 (n <= 4) packed one per byte (every slot
// is below 256), so a lane picks its tap's slot with one shift and mask.
__host__ __device__ constexpr uint32_t tap_bytes(int t0, int n) {
    uint32_t w = 0;
    for (int u = 0; u < n; ++u) w |= (uint32_t)tap_lds(t0 + u) << (8 * u);
    return w;
}
static_assert(tap_lds(26) < 256, "tap slots fit a byte");

struct ConvTiles {
    uint32_t ti, tj, tk;  // blocks per axis
};

inline ConvTiles conv_tiles(const Lvl& L) {
    return ConvTiles{(uint32_t)((L.nx + kTI - 1) / kTI), (uint32_t)((L.ny + kTJ - 1) / kTJ),
                     (uint32_t)((L.nz + kTK - 1) / kTK)};
}

// Inference-only extras of the first encoder convolutions.  FEAT: the input
// is not a tensor but the network's features of the residual, formed while
// staging the halo exactly as features_kernel forms them (r/rms on fluid
// cells, the one-hot class), so the feature tensor never goes to memory.
// POOL: the 2x2x2 average of the block's
// (bf16-rounded) outputs into the coarser grid, pool_kernel's float
// operations: the blocks are even-aligned, so every coarse cell's children
// lie in one block.
struct ConvExtra {
    const float* r = nullptr;
    const uint8_t* code = nullptr;
    const float* inv_rms = nullptr;  // 1/rms of r (rms_kernel)
    __bf16* pool_out = nullptr;
    int cnx = 0, cny = 0, cnz = 0;
};

// UPIN: `in` lives on the 2x-coarser grid (ny, nz = cny, cnz) and the
// convolution reads its nearest-neighbour upsampling, in[i/2][j/2][k/2], as
// up_kernel would have written it (the upsampled tensor never goes to
// memory).
template <int CIN, int COUT, bool UPIN = false, bool FEAT = false, bool POOL = false>
__global__ __launch_bounds__(256) void conv_mfma_kernel(
    Lvl L, ConvTiles T, const __bf16* __restrict__ in, const bf16x8* __restrict__ wfrag,
    const float* __restrict__ bias, int relu, __bf16* __restrict__ out, const __bf16* __restrict__ skip,
    __bf16* __restrict__ out_sum, const int* __restrict__ done, int cny = 0, int cnz = 0, ConvExtra X = ConvExtra{}) {
    if (done && *done) return;
    static_assert(CIN == 8 || CIN == 16 || CIN == 32, "8, 16 or 32 input channels");
    static_assert(!FEAT || CIN == 8, "the features are 8 channels");
    constexpr int KS = ksteps_of(CIN);
    constexpr int MT = mtiles_of(COUT);
    constexpr int NG = CIN / 8;  // channel groups
    __shared__ bf16x8 tile[NG * kHalo];
    __shared__ __align__(16) __bf16 otile[POOL ? kTI * kTJ * kTK * COUT : 8];
    const uint32_t nblk = gridDim.x;
    const uint32_t b = xcd_remap(blockIdx.x, nblk);
    const int tk = (int)(b % T.tk);
    const uint32_t r = b / T.tk;
    const int tj = (int)(r % T.tj), ti = (int)(r / T.tj);
    const int i0 = ti * kTI, j0 = tj * kTJ, k0 = tk * kTK;
    const int64_t plane = (int64_t)L.ny * L.nz;
    float inv = 0.0f;
    if constexpr (FEAT) inv = *X.inv_rms;

    // stage the halo block: thread t holds one (k, channel group) slot of a
    // halo row, (hk, q) = (col / NG, col % NG) with col = t mod (kHK NG), so
    // consecutive threads read consecutive 16-byte pieces of one row, and
    // walks the kHI x kHJ halo rows R at a time from row t / (kHK NG)
    // (threads past R kHK NG stay idle); the per-row work is one small
    // division.
    {
        constexpr int COLS = kHK * NG, R = 256 / COLS, ROWS = kHI * kHJ;
        const int col = (int)threadIdx.x % COLS, row0 = (int)threadIdx.x / COLS;
        const int hk = col / NG, q = col - (col / NG) * NG;
        const int gk = k0 - 1 + hk;
        const bool k_in = row0 < R && (unsigned)gk < (unsigned)L.nz;
#pragma unroll
        for (int ps = 0; ps < (ROWS + R - 1) / R; ++ps) {
            const int hr = row0 + ps * R;
            if (row0 >= R || hr >= ROWS) break;
            const int hi = hr / kHJ, hj = hr - (hr / kHJ) * kHJ;
            const int gi = i0 - 1 + hi, gj = j0 - 1 + hj;
            bf16x8 val = bf16x8{};
            if (k_in && (unsigned)gi < (unsigned)L.nx && (unsigned)gj < (unsigned)L.ny) {
                if constexpr (FEAT) {
                    const int64_t cell = (int64_t)gi * plane + (int64_t)gj * L.nz + gk;
                    const uint8_t cc = X.code[cell];
                    __bf16 f[8];
                    f[0] = (__bf16)(cc >= kCodeFluid0 ? X.r[cell] * inv : 0.0f);
                    f[1] = (__bf16)(cc >= kCodeFluid0 ? 1.0f : 0.0f);
                    f[2] = (__bf16)(cc == kCodeAir ? 1.0f : 0.0f);
                    f[3] = (__bf16)(cc == kCodeSolid ? 1.0f : 0.0f);
                    f[4] = f[5] = f[6] = f[7] = (__bf16)0.0f;
                    val = *reinterpret_cast<const bf16x8*>(f);
                } else {
                    const int64_t src = UPIN ? ((int64_t)(gi >> 1) * cny + (gj >> 1)) * cnz + (gk >> 1)
                                             : (int64_t)gi * plane + (int64_t)gj * L.nz + gk;
                    val = *reinterpret_cast<const bf16x8*>(in + src * CIN + q * 8);
                }
            }
            tile[q * kHalo + hr * kHK + hk] = val;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    f32x4 acc[MT][kTJ];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < kTJ; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Fully unrolled, so each k-step's taps are constants: lane group g of
    // k-step s takes channels 32 s + 8 g .. + 7 of the K axis, i.e. channel
    // group q of tap t = (32 s + 8 g) / CIN (t = 27 on the padding of the
    // last k-step, where B = 0).  Per step that is a select among at most four
    // constant LDS offsets instead of the tap's division and decomposition.
    const int lane_org = w * kHJ * kHK + c;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bf16x8 a[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) a[mt] = wfrag[((size_t)s * MT + mt) * 64 + lane];
        int off;
        bool live;
        if constexpr (CIN == 8) {  // t = 4 s + g, q = 0
            off = (int)((tap_bytes(4 * s, 4) >> (8 * g)) & 0xFFu);
            live = 4 * s + g < 27;
        } else if constexpr (CIN == 16) {  // t = 2 s + g / 2, q = g % 2
            off = (g & 1) * kHalo + (int)((tap_bytes(2 * s, 2) >> (8 * (g >> 1))) & 0xFFu);
            live = 2 * s + (g >> 1) < 27;
        } else {  // t = s, q = g
            off = g * kHalo + tap_lds(s);
            live = true;
        }
        const int hbase = lane_org + off;
#pragma unroll
        for (int nt = 0; nt < kTJ; ++nt) {
            const bf16x8 bv = live ? tile[hbase + nt * kHK] : bf16x8{};
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], bv, acc[mt][nt], 0, 0, 0);
        }
    }

    // Epilogue: lane holds rows (channels) 4g..4g+3 of tile mt for cell c of row nt.
    const int ci = i0 + w, ck = k0 + c;
    const bool ik_in = ci < L.nx && ck < L.nz;
    if (!POOL && !ik_in) return;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        if (!ik_in) break;
        const int co0 = mt * 16 + g * 4;
        if (co0 >= COUT) continue;
        float bv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = bias ? bias[co0 + q] : 0.0f;
#pragma unroll
        for (int nt = 0; nt < kTJ; ++nt) {
            const int cj = j0 + nt;
            if (cj >= L.ny) continue;
            __bf16 o4[4], s4[4];
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                v[q] = acc[mt][nt][q] + bv[q];
                if (relu) v[q] = v[q] > 0.0f ? v[q] : 0.0f;
                o4[q] = (__bf16)v[q];
            }
            const int64_t o = ((int64_t)ci * plane + (int64_t)cj * L.nz + ck) * COUT + co0;
            if (out) *reinterpret_cast<uint2*>(out + o) = *reinterpret_cast<const uint2*>(o4);
            if constexpr (POOL)
                *reinterpret_cast<uint2*>(otile + ((w * kTJ + nt) * kTK + c) * COUT + co0) =
                    *reinterpret_cast<const uint2*>(o4);
            if (skip) {
                const uint2 sk = *reinterpret_cast<const uint2*>(skip + o);
                const __bf16* skv = reinterpret_cast<const __bf16*>(&sk);
#pragma unroll
                for (int q = 0; q < 4; ++q) s4[q] = (__bf16)((float)o4[q] + (float)skv[q]);
                *reinterpret_cast<uint2*>(out_sum + o) = *reinterpret_cast<const uint2*>(s4);
            }
        }
    }
    if constexpr (POOL) {
        __syncthreads();
        // 2 x 2 x 8 coarse cells under the block, 8-channel groups
        constexpr int G8 = COUT / 8;
        const int t = threadIdx.x;
        if (t < 32 * G8) {
            const int lc = t / G8, gq = t - (t / G8) * G8;
            const int li = lc >> 4, lj = (lc >> 3) & 1, lk = lc & 7;
            const int I = i0 / 2 + li, J = j0 / 2 + lj, K = k0 / 2 + lk;
            if (I < X.cnx && J < X.cny && K < X.cnz) {
                float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                int cnt = 0;
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                        for (int d = 0; d < 2; ++d) {
                            const int fi = 2 * li + a, fj = 2 * lj + bb, fk = 2 * lk + d;
                            if (i0 + fi >= L.nx || j0 + fj >= L.ny || k0 + fk >= L.nz) continue;
                            const bf16x8 x =
                                *reinterpret_cast<const bf16x8*>(otile + ((fi * kTJ + fj) * kTK + fk) * COUT + gq * 8);
#pragma unroll
                            for (int q = 0; q < 8; ++q) acc[q] += (float)x[q];
                            ++cnt;
                        }
                const float pinv = 1.0f / (float)cnt;
                __bf16 o[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) o[q] = (__bf16)(acc[q] * pinv);
                const int64_t cc = ((int64_t)I * X.cny + J) * X.cnz + K;
                *reinterpret_cast<uint4*>(X.pool_out + cc * COUT + gq * 8) = *reinterpret_cast<const uint4*>(o);
            }
        }
    }
}

// Final 8 -> 1 convolution (VALU, FP32): y = b + the taps inside the grid, in
// tap order, each the 8 fmaf of its channels in channel order; z = scale * y
// on fluid cells, 0 elsewhere.  A workgroup owns one 4 x 4 x 16 block of the
// convolutions' tiling and stages its input with the one-cell halo in LDS
// (the layout of conv_mfma_kernel with one channel group), so the 27 taps of
// a cell are read from LDS instead of 27 gathers from L1/L2; thread (w, row,
// c) computes cell (i0 + w, j0 + row, k0 + c).
__global__ __launch_bounds__(256) void conv_out_kernel(Lvl L, ConvTiles T, const __bf16* __restrict__ in,
                                                       const float* __restrict__ w, const float* __restrict__ bias,
                                                       const float* __restrict__ scale_ptr,
                                                       const uint8_t* __restrict__ code,
                                                       float* __restrict__ y_out, float* __restrict__ z_out,
                                                       const int* __restrict__ done) {
    if (done && *done) return;
    __shared__ bf16x8 tile[kHalo];
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int tk = (int)(b % T.tk);
    const uint32_t rr = b / T.tk;
    const int tj = (int)(rr % T.tj), ti = (int)(rr / T.tj);
    const int i0 = ti * kTI, j0 = tj * kTJ, k0 = tk * kTK;
    const int64_t plane = (int64_t)L.ny * L.nz;
    {  // conv_mfma_kernel's staging with one channel group
        constexpr int R = 256 / kHK, ROWS = kHI * kHJ;
        const int hk = (int)threadIdx.x % kHK, row0 = (int)threadIdx.x / kHK;
        const int gk = k0 - 1 + hk;
        const bool k_in = row0 < R && (unsigned)gk < (unsigned)L.nz;
#pragma unroll
        for (int ps = 0; ps < (ROWS + R - 1) / R; ++ps) {
            const int hr = row0 + ps * R;
            if (row0 >= R || hr >= ROWS) break;
            const int hi = hr / kHJ, hj = hr - (hr / kHJ) * kHJ;
            const int gi = i0 - 1 + hi, gj = j0 - 1 + hj;
            bf16x8 val = bf16x8{};
            if (k_in && (unsigned)gi < (unsigned)L.nx && (unsigned)gj < (unsigned)L.ny)
                val = *reinterpret_cast<const bf16x8*>(in + ((int64_t)gi * plane + (int64_t)gj * L.nz + gk) * 8);
            tile[hr * kHK + hk] = val;
        }
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane >> 4, c = lane & 15;
    const int i = i0 + wv, j = j0 + row, k = k0 + c;
    if (i >= L.nx || j >= L.ny || k >= L.nz) return;
    const float scale = scale_ptr ? *scale_ptr : 1.0f;
    float acc = bias[0];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
        if ((unsigned)(i + di) >= (unsigned)L.nx || (unsigned)(j + dj) >= (unsigned)L.ny ||
            (unsigned)(k + dk) >= (unsigned)L.nz)
            continue;
        const bf16x8 x = tile[((wv + 1 + di) * kHJ + row + 1 + dj) * kHK + c + 1 + dk];
#pragma unroll
        for (int ci = 0; ci < 8; ++ci) acc = fmaf((float)x[ci], w[t * 8 + ci], acc);
    }
    const int64_t cell = (int64_t)i * plane + (int64_t)j * L.nz + k;
    if (y_out) y_out[cell] = acc;
    if (z_out) z_out[cell] = code[cell] >= kCodeFluid0 ? acc * scale : 0.0f;
}

// Input gradient of the final conv: g_d1[cell][ci] = sum_t gy[cell - off(t)] W[t][ci].
__global__ __launch_bounds__(256) void conv_out_bwd_data_kernel(Lvl L, const float* __restrict__ gy,
                                                                const float* __restrict__ w,
                                                                __bf16* __restrict__ gin) {
    const int64_t plane = (int64_t)L.ny * L.nz;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x; cell < L.n; cell += stride) {
        int i, j, k;
        cell_ijk(L, (uint32_t)cell, i, j, k);
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            // out cell o = cell - off(t) used in[cell] through tap t.
            if ((unsigned)(i - di) >= (unsigned)L.nx || (unsigned)(j - dj) >= (unsigned)L.ny ||
                (unsigned)(k - dk) >= (unsigned)L.nz)
                continue;
            const float gv = gy[cell - di * plane - dj * L.nz - dk];
#pragma unroll
            for (int ci = 0; ci < 8; ++ci) acc[ci] = fmaf(gv, w[t * 8 + ci], acc[ci]);
        }
        __bf16 o[8];
#pragma unroll
        for (int ci = 0; ci < 8; ++ci) o[ci] = (__bf16)acc[ci];
        *reinterpret_cast<uint4*>(gin + cell * 8) = *reinterpret_cast<const uint4*>(o);
    }
}

// Weight (and bias) gradient partials of one conv:
//   dW[t][ci][co] = sum_cells in[nbr(cell,t)][ci] * g[cell][co],  db[co] = sum g.
// A workgroup owns a contiguous cell range; thread p owns (t, ci) pairs
// p, p+256, ... and keeps COUT accumulators for each.  Partials go to
// part[wg][27*CIN*COUT + COUT] and are summed in workgroup order by
// wgrad_reduce_kernel.  gy (FP32, one channel) replaces g for the final conv.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void wgrad_kernel(Lvl L, const __bf16* __restrict__ in,
                                                    const __bf16* __restrict__ g, const float* __restrict__ gy,
                                                    float* __restrict__ part, int cells_per_wg) {
    constexpr int NPAIR = 27 * CIN;
    constexpr int PPT = (NPAIR + 255) / 256;  // pairs per thread
    constexpr int NOUT = NPAIR * COUT + COUT;
    const int64_t c0 = (int64_t)blockIdx.x * cells_per_wg;
    int64_t c1 = c0 + cells_per_wg;
    if (c1 > L.n) c1 = L.n;
    const int64_t plane = (int64_t)L.ny * L.nz;
    float acc[PPT][COUT];
#pragma unroll
    for (int q = 0; q < PPT; ++q)
#pragma unroll
        for (int co = 0; co < COUT; ++co) acc[q][co] = 0.0f;
    float bacc = 0.0f;  // bias gradient: thread co < COUT
    for (int64_t cell = c0; cell < c1; ++cell) {
        int i, j, k;
        cell_ijk(L, (uint32_t)cell, i, j, k);
        float gv[COUT];
        if constexpr (COUT == 1) {
            gv[0] = gy[cell];
        } else {
#pragma unroll
            for (int co = 0; co < COUT; ++co) gv[co] = (float)g[cell * COUT + co];
        }
#pragma unroll
        for (int co = 0; co < COUT; ++co)
            if ((int)threadIdx.x == co) bacc += gv[co];
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * 256;
            if (p >= NPAIR) break;
            const int t = p / CIN, ci = p - t * CIN;
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            if ((unsigned)(i + di) >= (unsigned)L.nx || (unsigned)(j + dj) >= (unsigned)L.ny ||
                (unsigned)(k + dk) >= (unsigned)L.nz)
                continue;
            const float xv = (float)in[(cell + di * plane + dj * L.nz + dk) * CIN + ci];
#pragma unroll
            for (int co = 0; co < COUT; ++co) acc[q][co] = fmaf(xv, gv[co], acc[q][co]);
        }
    }
    float* dst = part + (size_t)blockIdx.x * NOUT;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int p = threadIdx.x + q * 256;
        if (p >= NPAIR) break;
#pragma unroll
        for (int co = 0; co < COUT; ++co) dst[(size_t)p * COUT + co] = acc[q][co];
    }
    if (threadIdx.x < COUT) dst[NPAIR * COUT + threadIdx.x] = bacc;
}

// grad[o] (+)= sum over workgroups of part[wg][o], in workgroup order.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ part, int nwg, int nout,
                                                           float* __restrict__ grad, int accumulate) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= nout) return;
    double s = 0.0;
    for (int w = 0; w < nwg; ++w) s += (double)part[(size_t)w * nout + o];
    grad[o] = accumulate ? grad[o] + (float)s : (float)s;
}

// ---------------------------------------------------------------------------
// Elementwise / resampling kernels (channels-last BF16)
// ---------------------------------------------------------------------------

// rms = sqrt(r.r / n_fluid) from the r.r partials (1 when partials is null)
// into rms[0], and its reciprocal (0 when rms is 0) into rms[1].  One
// workgroup, so the network's kernels read the two values instead of each
// summing the partials again.
__global__ __launch_bounds__(256) void rms_kernel(const double* __restrict__ rr_partials, unsigned G, double n_fluid,
                                                  const unsigned* __restrict__ n_fluid_dev, float* __restrict__ rms,
                                                  const int* __restrict__ done) {
    if (done && *done) return;
    double v[1] = {0.0};
    if (rr_partials) sum_partials<1>(rr_partials, G, v);
    const double nf = n_fluid_dev ? (double)*n_fluid_dev : n_fluid;
    const double rv = rr_partials ? sqrt(v[0] / (nf > 0 ? nf : 1.0)) : 1.0;
    if (threadIdx.x == 0) {
        rms[0] = (float)rv;
        rms[1] = rv > 0.0 ? (float)(1.0 / rv) : 0.0f;
    }
}

// Features of the residual: r/rms (inv_rms = 1/rms from rms_kernel), one-hot class.
__global__ __launch_bounds__(256) void features_kernel(Lvl L, const float* __restrict__ r,
                                                       const uint8_t* __restrict__ code,
                                                       const float* __restrict__ inv_rms,
                                                       __bf16* __restrict__ x0, const int* __restrict__ done) {
    if (done && *done) return;
    const float inv = *inv_rms;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x; cell < L.n; cell += stride) {
        const uint8_t cc = code[cell];
        __bf16 f[8];
        f[0] = (__bf16)(cc >= kCodeFluid0 ? r[cell] * inv : 0.0f);
        f[1] = (__bf16)(cc >= kCodeFluid0 ? 1.0f : 0.0f);
        f[2] = (__bf16)(cc == kCodeAir ? 1.0f : 0.0f);
        f[3] = (__bf16)(cc == kCodeSolid ? 1.0f : 0.0f);
        f[4] = f[5] = f[6] = f[7] = (__bf16)0.0f;
        *reinterpret_cast<uint4*>(x0 + cell * 8) = *reinterpret_cast<const uint4*>(f);
    }
}

// r.r partials over fluid cells.
__global__ __launch_bounds__(256) void rr_partials_kernel(int64_t n, const float* __restrict__ r,
                                                          double* __restrict__ partials) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) s += (double)r[t] * (double)r[t];
    double v[1] = {s};
    write_partials<1>(v, partials);
}

// Average of the existing children (2x2x2, ceil sizes); C channels.
template <int C>
__global__ __launch_bounds__(256) void pool_kernel(Lvl F, Lvl Cg, const __bf16* __restrict__ fin,
                                                   __bf16* __restrict__ cout_, const int* __restrict__ done) {
    if (done && *done) return;
    constexpr int G8 = C / 8;  // 8-channel groups
    const int64_t total = Cg.n * G8;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
        const int64_t cc = t / G8;
        const int gi = (int)(t - cc * G8);
        int I, J, K;
        cell_ijk(Cg, (uint32_t)cc, I, J, K);
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int cnt = 0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + d;
                    if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
                    const int64_t fc = ((int64_t)i * F.ny + j) * F.nz + k;
                    const bf16x8 x = *reinterpret_cast<const bf16x8*>(fin + fc * C + gi * 8);
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[q] += (float)x[q];
                    ++cnt;
                }
        const float inv = 1.0f / (float)cnt;
        __bf16 o[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) o[q] = (__bf16)(acc[q] * inv);
        *reinterpret_cast<uint4*>(cout_ + cc * C + gi * 8) = *reinterpret_cast<const uint4*>(o);
    }
}

// Gradient of pool: g_fine[child] = g_coarse[parent] / count(parent)  (+= if accumulate).
template <int C>
__global__ __launch_bounds__(256) void pool_bwd_kernel(Lvl F, Lvl Cg, const __bf16* __restrict__ gc,
                                                       __bf16* __restrict__ gf, int accumulate) {
    constexpr int G8 = C / 8;
    const int64_t total = F.n * G8;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
        const int64_t fc = t / G8;
        const int gi = (int)(t - fc * G8);
        int i, j, k;
        cell_ijk(F, (uint32_t)fc, i, j, k);
        const int I = i >> 1, J = j >> 1, K = k >> 1;
        const int cnt = (2 * I + 1 < F.nx ? 2 : 1) * (2 * J + 1 < F.ny ? 2 : 1) * (2 * K + 1 < F.nz ? 2 : 1);
        const int64_t cc = ((int64_t)I * Cg.ny + J) * Cg.nz + K;
        const bf16x8 x = *reinterpret_cast<const bf16x8*>(gc + cc * C + gi * 8);
        __bf16 o[8];
        const float inv = 1.0f / (float)cnt;
        bf16x8 prev{};
        if (accumulate) prev = *reinterpret_cast<const bf16x8*>(gf + fc * C + gi * 8);
#pragma unroll
        for (int q = 0; q < 8; ++q) o[q] = (__bf16)((float)x[q] * inv + (accumulate ? (float)prev[q] : 0.0f));
        *reinterpret_cast<uint4*>(gf + fc * C + gi * 8) = *reinterpret_cast<const uint4*>(o);
    }
}

// Nearest-neighbour upsampling: fine[child] = coarse[parent].
template <int C>
__global__ __launch_bounds__(256) void up_kernel(Lvl F, Lvl Cg, const __bf16* __restrict__ cin_,
                                                 __bf16* __restrict__ fout, const int* __restrict__ done) {
    if (done && *done) return;
    constexpr int G8 = C / 8;
    const int64_t total = F.n * G8;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
        const int64_t fc = t / G8;
        const int gi = (int)(t - fc * G8);
        int i, j, k;
        cell_ijk(F, (uint32_t)fc, i, j, k);
        const int64_t cc = ((int64_t)(i >> 1) * Cg.ny + (j >> 1)) * Cg.nz + (k >> 1);
        *reinterpret_cast<uint4*>(fout + fc * C + gi * 8) =
            *reinterpret_cast<const uint4*>(cin_ + cc * C + gi * 8);
    }
}

// Gradient of up: g_coarse[parent] = sum over its children of g_fine (+= skip grad when given).
template <int C>
__global__ __launch_bounds__(256) void up_bwd_kernel(Lvl F, Lvl Cg, const __bf16* __restrict__ gf,
                                                     const __bf16* __restrict__ gadd, __bf16* __restrict__ gc) {
    constexpr int G8 = C / 8;
    const int64_t total = Cg.n * G8;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
        const int64_t cc = t / G8;
        const int gi = (int)(t - cc * G8);
        int I, J, K;
        cell_ijk(Cg, (uint32_t)cc, I, J, K);
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (gadd) {
            const bf16x8 x = *reinterpret_cast<const bf16x8*>(gadd + cc * C + gi * 8);
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] = (float)x[q];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + d;
                    if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
                    const int64_t fc = ((int64_t)i * F.ny + j) * F.nz + k;
                    const bf16x8 x = *reinterpret_cast<const bf16x8*>(gf + fc * C + gi * 8);
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[q] += (float)x[q];
                }
        __bf16 o[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) o[q] = (__bf16)acc[q];
        *reinterpret_cast<uint4*>(gc + cc * C + gi * 8) = *reinterpret_cast<const uint4*>(o);
    }
}

// ReLU gradient: g *= (act > 0), optionally adding a second gradient first.
__global__ __launch_bounds__(256) void relu_bwd_kernel(int64_t nelem, const __bf16* __restrict__ act,
                                                       const __bf16* __restrict__ g_in,
                                                       const __bf16* __restrict__ g_add, __bf16* __restrict__ g_out) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nelem; t += stride) {
        float gv = (float)g_in[t];
        if (g_add) gv += (float)g_add[t];
        g_out[t] = (__bf16)((float)act[t] > 0.0f ? gv : 0.0f);
    }
}

// Packs FP32 weights into per-lane MFMA A fragments.
//   fwd: row co, k = t*cin + ci  -> W[t][ci][co]
//   bwd: row ci, k = t'*cout' + co (cout' = cout padded to >= 8), t' = 26 - t
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ params, __bf16* __restrict__ packed) {
    for (int l = 0; l < kNumConv; ++l) {
        const ConvDef d = conv_def(l);
        const float* W = params + conv_w_off(l);
        for (int bwd = 0; bwd < 2; ++bwd) {
            if (l == kNumConv - 1) continue;  // the final conv runs on VALU both ways
            const int kin = bwd ? d.cout : d.cin;     // GEMM K channels
            const int rows = bwd ? d.cin : d.cout;    // GEMM M rows
            const int KS = ksteps_of(kin), MT = mtiles_of(rows);
            __bf16* dst = packed + frag_off(l, bwd);
            const int total = KS * MT * 64 * 8;
            for (int e = blockIdx.x * 256 + threadIdx.x; e < total; e += gridDim.x * 256) {
                const int q = e & 7;
                const int lane = (e >> 3) & 63;
                const int mt = (e >> 9) % MT;
                const int s = (e >> 9) / MT;
                const int row = mt * 16 + (lane & 15);
                const int k = s * 32 + (lane >> 4) * 8 + q;
                const int t = k / kin, ch = k - t * kin;
                float v = 0.0f;
                if (t < 27 && row < rows) {
                    if (!bwd) v = W[((size_t)t * d.cin + ch) * d.cout + row];
                    else v = W[((size_t)(26 - t) * d.cin + row) * d.cout + ch];
                }
                dst[e] = (__bf16)v;
            }
        }
    }
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2,
                                                   float gscale) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float gr = g[t] * gscale;
    const float mm = b1 * m[t] + (1.0f - b1) * gr;
    const float vv = b2 * v[t] + (1.0f - b2) * gr * gr;
    m[t] = mm;
    v[t] = vv;
    w[t] = w[t] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// d = z - e (FP32), the error of the network's correction.
__global__ __launch_bounds__(256) void sub_kernel(int64_t n, const float* __restrict__ z,
                                                  const float* __restrict__ e, float* __restrict__ d) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) d[t] = z[t] - e[t];
}

// dL/dy for the A-norm loss L = (z-e)^T A (z-e) / (e.r) (e.r = e^T A e for
// the exact correction), z = scale*y on fluid cells:
// dL/dy = 2*scale*A(z-e)/(e.r) on fluid cells.  ad = A(z-e), d = z-e,
// er_ptr -> e.r (device).  Also partials of d.A d (the loss numerator).
__global__ __launch_bounds__(256) void loss_grad_kernel(int64_t n, const float* __restrict__ ad,
                                                        const float* __restrict__ d,
                                                        const uint8_t* __restrict__ code,
                                                        const float* __restrict__ scale_ptr, const double* er_ptr,
                                                        float* __restrict__ gy, double* __restrict__ partials) {
    const double er = *er_ptr;
    const float sc = *scale_ptr;
    const float f = er != 0.0 ? (float)(2.0 / er) : 0.0f;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) {
        const bool fl = code[t] >= kCodeFluid0;
        gy[t] = fl ? f * sc * ad[t] : 0.0f;
        if (fl) acc += (double)d[t] * (double)ad[t];
    }
    double v[1] = {acc};
    write_partials<1>(v, partials);
}

// e.r over fluid cells: partials then totals (loss normaliser).
__global__ __launch_bounds__(256) void dot_partials_kernel(int64_t n, const float* __restrict__ a,
                                                           const float* __restrict__ b,
                                                           double* __restrict__ partials) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) s += (double)a[t] * (double)b[t];
    double v[1] = {s};
    write_partials<1>(v, partials);
}

__global__ __launch_bounds__(256) void dot_total_kernel(const double* __restrict__ partials, unsigned G,
                                                        double* __restrict__ out) {
    double v[1];
    sum_partials<1>(partials, G, v);
    if (threadIdx.x == 0) *out = v[0];
}

}  // namespace bs

// ===========================================================================
// Host side
// ===========================================================================
namespace bs {

// Ap with class codes (the PCG operator), for the loss.
__global__ __launch_bounds__(256) void code_spmv_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                        const float* __restrict__ p, float* __restrict__ ap) {
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<1>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        Nbhd<1> nb;
        nb.load(T, code, p, i, j, k0, c0);
        ap[c0] = nb.apply(0);
    }
}

// Class codes from the int32 mask (air 0, solid 1, fluid 2 + #solid
// neighbours; out of bounds = air).
__global__ __launch_bounds__(256) void build_codes_kernel(Lvl L, const int* __restrict__ mat, int mf, int ma,
                                                          uint8_t* __restrict__ code) {
    const int64_t plane = (int64_t)L.ny * L.nz;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < L.n; c += stride) {
        const int m = mat[c];
        if (m != mf) { code[c] = m == ma ? kCodeAir : kCodeSolid; continue; }
        int i, j, k;
        cell_ijk(L, (uint32_t)c, i, j, k);
        int ns = 0;
        auto solid = [&](int64_t o) { const int q = mat[o]; return (q != mf && q != ma) ? 1 : 0; };
        if (i > 0) ns += solid(c - plane);
        if (i < L.nx - 1) ns += solid(c + plane);
        if (j > 0) ns += solid(c - L.nz);
        if (j < L.ny - 1) ns += solid(c + L.nz);
        if (k > 0) ns += solid(c - 1);
        if (k < L.nz - 1) ns += solid(c + 1);
        code[c] = (uint8_t)(kCodeFluid0 + ns);
    }
}

static inline size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

struct UNetWs {
    Lvl L1, L2, L3;
    // forward
    __bf16 *x0, *h1, *p1, *h2, *p2, *h3a, *h3, *up3, *u2, *s2, *up2, *u1, *s1, *d1;
    float *y, *rms;
    double* part;  // reduction partials (2 * kPartWg)
    // backward (training only)
    float *gy, *az, *wpart;
    __bf16 *g_d1p, *g_s1, *g_u1p, *g_up2, *g_s2, *g_u2p, *g_up3, *g_h3, *g_h3p, *g_h3a, *g_h3ap, *g_p2,
        *g_h2, *g_h2p, *g_p1, *g_h1, *g_h1p;
    double* scal;  // [4]: e.r, loss partial total, ...
    size_t bytes;
};

constexpr unsigned kPartWg = 1024;
constexpr int kWgradMaxWg = 1024;
constexpr size_t kWgradMaxOut = 27 * 32 * 32 + 32;

static UNetWs carve_unet(void* base, int nx, int ny, int nz, int training) {
    UNetWs W;
    memset(&W, 0, sizeof(W));
    W.L1 = make_lvl(nx, ny, nz);
    W.L2 = make_lvl((nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2);
    W.L3 = make_lvl((W.L2.nx + 1) / 2, (W.L2.ny + 1) / 2, (W.L2.nz + 1) / 2);
    const size_t n1 = W.L1.n, n2 = W.L2.n, n3 = W.L3.n;
    char* c = (char*)base;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? c + off : nullptr;
        off += al(bytes);
        return p;
    };
    const size_t B = sizeof(__bf16);
    W.x0 = (__bf16*)take(n1 * 8 * B);
    W.h1 = (__bf16*)take(n1 * 8 * B);
    W.p1 = (__bf16*)take(n2 * 8 * B);
    W.h2 = (__bf16*)take(n2 * 16 * B);
    W.p2 = (__bf16*)take(n3 * 16 * B);
    W.h3a = (__bf16*)take(n3 * 32 * B);
    W.h3 = (__bf16*)take(n3 * 32 * B);
    W.up3 = (__bf16*)take(n2 * 32 * B);
    W.s2 = (__bf16*)take(n2 * 16 * B);
    W.up2 = (__bf16*)take(n1 * 16 * B);
    W.s1 = (__bf16*)take(n1 * 8 * B);
    W.d1 = (__bf16*)take(n1 * 8 * B);
    W.y = (float*)take(n1 * 4);
    W.rms = (float*)take(64);
    W.part = (double*)take(2 * kPartWg * 8);
    if (training) {
        W.u2 = (__bf16*)take(n2 * 16 * B);
        W.u1 = (__bf16*)take(n1 * 8 * B);
        W.gy = (float*)take(n1 * 4);
        W.az = (float*)take(n1 * 4);
        W.wpart = (float*)take((size_t)kWgradMaxWg * kWgradMaxOut * 4);
        W.g_d1p = (__bf16*)take(n1 * 8 * B);
        W.g_s1 = (__bf16*)take(n1 * 8 * B);
        W.g_u1p = (__bf16*)take(n1 * 8 * B);
        W.g_up2 = (__bf16*)take(n1 * 16 * B);
        W.g_s2 = (__bf16*)take(n2 * 16 * B);
        W.g_u2p = (__bf16*)take(n2 * 16 * B);
        W.g_up3 = (__bf16*)take(n2 * 32 * B);
        W.g_h3 = (__bf16*)take(n3 * 32 * B);
        W.g_h3p = (__bf16*)take(n3 * 32 * B);
        W.g_h3a = (__bf16*)take(n3 * 32 * B);
        W.g_h3ap = (__bf16*)take(n3 * 32 * B);
        W.g_p2 = (__bf16*)take(n3 * 16 * B);
        W.g_h2 = (__bf16*)take(n2 * 16 * B);
        W.g_h2p = (__bf16*)take(n2 * 16 * B);
        W.g_p1 = (__bf16*)take(n2 * 8 * B);
        W.g_h1 = (__bf16*)take(n1 * 8 * B);
        W.g_h1p = (__bf16*)take(n1 * 8 * B);
        W.scal = (double*)take(64);
    }
    W.bytes = off;
    return W;
}

static unsigned elem_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

// up_from: `in` is on that (2x-coarser) level and is read upsampled.
template <int CIN, int COUT>
static void conv_fwd(hipStream_t st, const Lvl& L, const __bf16* in, const __bf16* packed, int l,
                     const float* params, int relu, __bf16* out, const __bf16* skip, __bf16* out_sum,
                     const int* done, const Lvl* up_from = nullptr) {
    const ConvTiles T = conv_tiles(L);
    const bf16x8* wf = reinterpret_cast<const bf16x8*>(packed + frag_off(l, 0));
    if (up_from)
        conv_mfma_kernel<CIN, COUT, true><<<T.ti * T.tj * T.tk, 256, 0, st>>>(
            L, T, in, wf, params + conv_b_off(l), relu, out, skip, out_sum, done, up_from->ny, up_from->nz);
    else
        conv_mfma_kernel<CIN, COUT><<<T.ti * T.tj * T.tk, 256, 0, st>>>(L, T, in, wf, params + conv_b_off(l),
                                                                        relu, out, skip, out_sum, done);
}

// conv_fwd with the inference extras (ConvExtra: features formed while
// staging, pooled output).
template <int CIN, int COUT, bool FEAT, bool POOL>
static void conv_fwd_x(hipStream_t st, const Lvl& L, const __bf16* in, const __bf16* packed, int l,
                       const float* params, int relu, __bf16* out, const int* done, const ConvExtra& X) {
    const ConvTiles T = conv_tiles(L);
    const bf16x8* wf = reinterpret_cast<const bf16x8*>(packed + frag_off(l, 0));
    conv_mfma_kernel<CIN, COUT, false, FEAT, POOL><<<T.ti * T.tj * T.tk, 256, 0, st>>>(
        L, T, in, wf, params + conv_b_off(l), relu, out, nullptr, nullptr, done, 0, 0, X);
}

// Input gradient of conv l: the forward kernel on g with the flipped weights.
template <int GC, int OC>
static void conv_bwd_data(hipStream_t st, const Lvl& L, const __bf16* g, const __bf16* packed, int l,
                          __bf16* gin) {
    const ConvTiles T = conv_tiles(L);
    conv_mfma_kernel<GC, OC><<<T.ti * T.tj * T.tk, 256, 0, st>>>(
        L, T, g, reinterpret_cast<const bf16x8*>(packed + frag_off(l, 1)), nullptr, 0, gin, nullptr, nullptr,
        nullptr);
}

template <int CIN, int COUT>
static void conv_wgrad(hipStream_t st, const Lvl& L, const __bf16* in, const __bf16* g, const float* gy,
                       float* wpart, int l, float* grads) {
    int cpw = (int)((L.n + kWgradMaxWg - 1) / kWgradMaxWg);
    if (cpw < 16) cpw = 16;
    const int nwg = (int)((L.n + cpw - 1) / cpw);
    wgrad_kernel<CIN, COUT><<<nwg, 256, 0, st>>>(L, in, g, gy, wpart, cpw);
    const int nout = 27 * CIN * COUT + COUT;
    wgrad_reduce_kernel<<<(nout + 255) / 256, 256, 0, st>>>(wpart, nwg, nout, grads + conv_w_off(l), 1);
}

// Forward pass; r -> z (FP32, fluid cells).  rr_partials/G: r.r partials
// already on the device (the Krylov loop has them); when null they are formed
// here.  The fluid count that normalises r is n_fluid_dev's value when that is
// not null (no host round trip), else n_fluid.  training != 0 also keeps the
// pre-skip activations for backward.
static int unet_forward(hipStream_t st, UNetWs& W, const float* params, const __bf16* packed, const float* r,
                        const uint8_t* code, double n_fluid, const unsigned* n_fluid_dev,
                        const double* rr_partials, unsigned G, float* z, const int* done, int training) {
    const unsigned eg1 = elem_grid(W.L1.n);
    if (!rr_partials) {
        rr_partials_kernel<<<kPartWg, 256, 0, st>>>(W.L1.n, r, W.part);
        rr_partials = W.part;
        G = kPartWg;
    }
    rms_kernel<<<1, 256, 0, st>>>(rr_partials, G, n_fluid, n_fluid_dev, W.rms, done);
    if (training) {
        features_kernel<<<eg1, 256, 0, st>>>(W.L1, r, code, W.rms + 1, W.x0, done);
        conv_fwd<8, 8>(st, W.L1, W.x0, packed, 0, params, 1, W.h1, nullptr, nullptr, done);
        pool_kernel<8><<<elem_grid(W.L2.n), 256, 0, st>>>(W.L1, W.L2, W.h1, W.p1, done);
        conv_fwd<8, 16>(st, W.L2, W.p1, packed, 1, params, 1, W.h2, nullptr, nullptr, done);
        pool_kernel<16><<<elem_grid(W.L3.n * 2), 256, 0, st>>>(W.L2, W.L3, W.h2, W.p2, done);
    } else {
        // the same values without the feature tensor and the two pooling
        // passes (GPU test: bit-identical to the training forward)
        ConvExtra X1;
        X1.r = r; X1.code = code; X1.inv_rms = W.rms + 1;
        X1.pool_out = W.p1; X1.cnx = W.L2.nx; X1.cny = W.L2.ny; X1.cnz = W.L2.nz;
        conv_fwd_x<8, 8, true, true>(st, W.L1, nullptr, packed, 0, params, 1, W.h1, done, X1);
        ConvExtra X2;
        X2.pool_out = W.p2; X2.cnx = W.L3.nx; X2.cny = W.L3.ny; X2.cnz = W.L3.nz;
        conv_fwd_x<8, 16, false, true>(st, W.L2, W.p1, packed, 1, params, 1, W.h2, done, X2);
    }
    conv_fwd<16, 32>(st, W.L3, W.p2, packed, 2, params, 1, W.h3a, nullptr, nullptr, done);
    conv_fwd<32, 32>(st, W.L3, W.h3a, packed, 3, params, 1, W.h3, nullptr, nullptr, done);
    // The decoder convolutions read their coarser input upsampled on the fly;
    // training keeps the upsampled tensors for the backward pass.
    if (training) {
        up_kernel<32><<<elem_grid(W.L2.n * 4), 256, 0, st>>>(W.L2, W.L3, W.h3, W.up3, done);
        conv_fwd<32, 16>(st, W.L2, W.up3, packed, 4, params, 1, W.u2, W.h2, W.s2, done);
        up_kernel<16><<<elem_grid(W.L1.n * 2), 256, 0, st>>>(W.L1, W.L2, W.s2, W.up2, done);
        conv_fwd<16, 8>(st, W.L1, W.up2, packed, 5, params, 1, W.u1, W.h1, W.s1, done);
    } else {
        conv_fwd<32, 16>(st, W.L2, W.h3, packed, 4, params, 1, nullptr, W.h2, W.s2, done, &W.L3);
        conv_fwd<16, 8>(st, W.L1, W.s2, packed, 5, params, 1, nullptr, W.h1, W.s1, done, &W.L2);
    }
    conv_fwd<8, 8>(st, W.L1, W.s1, packed, 6, params, 1, W.d1, nullptr, nullptr, done);
    const ConvTiles T1 = conv_tiles(W.L1);
    conv_out_kernel<<<T1.ti * T1.tj * T1.tk, 256, 0, st>>>(W.L1, T1, W.d1, params + conv_w_off(7),
                                                          params + conv_b_off(7), W.rms, code, W.y, z, done);
    return (int)hipGetLastError();
}

}  // namespace bs

using namespace bs;

extern "C" size_t bs_unet_num_params(void) { return kNumParams; }
extern "C" size_t bs_unet_packed_bytes(void) { return kPackedElems * sizeof(__bf16); }

extern "C" size_t bs_unet_workspace_bytes(int nx, int ny, int nz, int training) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return carve_unet(nullptr, nx, ny, nz, training).bytes;
}

extern "C" int bs_unet_pack(const float* params, void* packed, bs_stream_t stream) {
    pack_kernel<<<64, 256, 0, stream>>>(params, (__bf16*)packed);
    return (int)hipGetLastError();
}

extern "C" int bs_build_codes(const int* mat, int nx, int ny, int nz, int mat_fluid, int mat_air,
                              unsigned char* code, bs_stream_t stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    const Lvl L = make_lvl(nx, ny, nz);
    build_codes_kernel<<<elem_grid(L.n), 256, 0, stream>>>(L, mat, mat_fluid, mat_air, code);
    return (int)hipGetLastError();
}

extern "C" int bs_unet_forward(const float* params, const void* packed, const float* r,
                               const unsigned char* code, int nx, int ny, int nz, double n_fluid, float* z,
                               void* workspace, size_t workspace_bytes, int training, bs_stream_t stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    UNetWs W = carve_unet(workspace, nx, ny, nz, training);
    if (workspace_bytes < W.bytes) return (int)hipErrorInvalidValue;
    return unet_forward(stream, W, params, (const __bf16*)packed, r, code, n_fluid, nullptr, nullptr, 0, z, nullptr,
                        training);
}

// After bs_unet_forward(training=1) on (r, ...): loss and gradients for the
// target correction e (A e = r).  grads (FP32, bs_unet_num_params) are
// accumulated; loss_dev (2 device doubles) receives (z-e)^T A (z-e) and e.r,
// the loss being their ratio.
extern "C" int bs_unet_backward(const float* params, const void* packed_v, const float* r, const float* e,
                                const unsigned char* code, int nx, int ny, int nz, float* grads,
                                double* loss_dev, void* workspace, size_t workspace_bytes, bs_stream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    UNetWs W = carve_unet(workspace, nx, ny, nz, 1);
    if (workspace_bytes < W.bytes) return (int)hipErrorInvalidValue;
    const __bf16* packed = (const __bf16*)packed_v;
    const unsigned eg1 = elem_grid(W.L1.n);
    // z (recomputed from d1) -> d = z - e -> A d; e.r normaliser.
    const ConvTiles T1 = conv_tiles(W.L1);
    conv_out_kernel<<<T1.ti * T1.tj * T1.tk, 256, 0, st>>>(W.L1, T1, W.d1, params + conv_w_off(7),
                                                          params + conv_b_off(7), W.rms, code, nullptr, W.gy,
                                                          nullptr);
    sub_kernel<<<eg1, 256, 0, st>>>(W.L1.n, W.gy, e, W.gy);
    const Tiles T = make_tiles(nx, ny, nz, 1);
    const unsigned nwg = T.ntiles < 2048 ? T.ntiles : 2048;
    code_spmv_kernel<<<nwg, 256, 0, st>>>(T, code, W.gy, W.az);
    dot_partials_kernel<<<kPartWg, 256, 0, st>>>(W.L1.n, e, r, W.part);
    dot_total_kernel<<<1, 256, 0, st>>>(W.part, kPartWg, W.scal);
    // dL/dy into gy (overwrites d: each element is read before it is written).
    loss_grad_kernel<<<kPartWg, 256, 0, st>>>(W.L1.n, W.az, W.gy, code, W.rms, W.scal, W.gy, W.part);
    dot_total_kernel<<<1, 256, 0, st>>>(W.part, kPartWg, W.scal + 1);
    // c8 (final, VALU)
    conv_wgrad<8, 1>(st, W.L1, W.d1, nullptr, W.gy, W.wpart, 7, grads);
    conv_out_bwd_data_kernel<<<eg1, 256, 0, st>>>(W.L1, W.gy, params + conv_w_off(7), W.g_d1p);
    relu_bwd_kernel<<<elem_grid(W.L1.n * 8), 256, 0, st>>>(W.L1.n * 8, W.d1, W.g_d1p, nullptr, W.g_d1p);
    // c7: s1 -> d1
    conv_wgrad<8, 8>(st, W.L1, W.s1, W.g_d1p, nullptr, W.wpart, 6, grads);
    conv_bwd_data<8, 8>(st, W.L1, W.g_d1p, packed, 6, W.g_s1);
    // s1 = u1 + h1: u1 branch
    relu_bwd_kernel<<<elem_grid(W.L1.n * 8), 256, 0, st>>>(W.L1.n * 8, W.u1, W.g_s1, nullptr, W.g_u1p);
    // c6: up2 -> u1
    conv_wgrad<16, 8>(st, W.L1, W.up2, W.g_u1p, nullptr, W.wpart, 5, grads);
    conv_bwd_data<8, 16>(st, W.L1, W.g_u1p, packed, 5, W.g_up2);
    up_bwd_kernel<16><<<elem_grid(W.L2.n * 2), 256, 0, st>>>(W.L1, W.L2, W.g_up2, nullptr, W.g_s2);
    // s2 = u2 + h2: u2 branch
    relu_bwd_kernel<<<elem_grid(W.L2.n * 16), 256, 0, st>>>(W.L2.n * 16, W.u2, W.g_s2, nullptr, W.g_u2p);
    // c5: up3 -> u2
    conv_wgrad<32, 16>(st, W.L2, W.up3, W.g_u2p, nullptr, W.wpart, 4, grads);
    conv_bwd_data<16, 32>(st, W.L2, W.g_u2p, packed, 4, W.g_up3);
    up_bwd_kernel<32><<<elem_grid(W.L3.n * 4), 256, 0, st>>>(W.L2, W.L3, W.g_up3, nullptr, W.g_h3);
    relu_bwd_kernel<<<elem_grid(W.L3.n * 32), 256, 0, st>>>(W.L3.n * 32, W.h3, W.g_h3, nullptr, W.g_h3p);
    // c4: h3a -> h3
    conv_wgrad<32, 32>(st, W.L3, W.h3a, W.g_h3p, nullptr, W.wpart, 3, grads);
    conv_bwd_data<32, 32>(st, W.L3, W.g_h3p, packed, 3, W.g_h3a);
    relu_bwd_kernel<<<elem_grid(W.L3.n * 32), 256, 0, st>>>(W.L3.n * 32, W.h3a, W.g_h3a, nullptr, W.g_h3ap);
    // c3: p2 -> h3a
    conv_wgrad<16, 32>(st, W.L3, W.p2, W.g_h3ap, nullptr, W.wpart, 2, grads);
    conv_bwd_data<32, 16>(st, W.L3, W.g_h3ap, packed, 2, W.g_p2);
    pool_bwd_kernel<16><<<elem_grid(W.L2.n * 2), 256, 0, st>>>(W.L2, W.L3, W.g_p2, W.g_h2, 0);
    // h2 gradient = pool path + skip (g_s2)
    relu_bwd_kernel<<<elem_grid(W.L2.n * 16), 256, 0, st>>>(W.L2.n * 16, W.h2, W.g_h2, W.g_s2, W.g_h2p);
    // c2: p1 -> h2
    conv_wgrad<8, 16>(st, W.L2, W.p1, W.g_h2p, nullptr, W.wpart, 1, grads);
    conv_bwd_data<16, 8>(st, W.L2, W.g_h2p, packed, 1, W.g_p1);
    pool_bwd_kernel<8><<<elem_grid(W.L1.n), 256, 0, st>>>(W.L1, W.L2, W.g_p1, W.g_h1, 0);
    relu_bwd_kernel<<<elem_grid(W.L1.n * 8), 256, 0, st>>>(W.L1.n * 8, W.h1, W.g_h1, W.g_s1, W.g_h1p);
    // c1: x0 -> h1
    conv_wgrad<8, 8>(st, W.L1, W.x0, W.g_h1p, nullptr, W.wpart, 0, grads);
    if (loss_dev) {
        // loss_dev[0] = (z-e).A(z-e), loss_dev[1] = e.r; L = [0] / [1].
        BS_RETURN_IF_ERR(hipMemcpyAsync(loss_dev, W.scal + 1, sizeof(double), hipMemcpyDeviceToDevice, st));
        BS_RETURN_IF_ERR(hipMemcpyAsync(loss_dev + 1, W.scal, sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    return (int)hipGetLastError();
}

extern "C" int bs_unet_adam(float* params, const float* grads, float* m, float* v, int step, float lr,
                            float gscale, bs_stream_t stream) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    const int n = (int)kNumParams;
    adam_kernel<<<(n + 255) / 256, 256, 0, stream>>>(params, grads, m, v, n, lr, b1, b2, eps, bc1, bc2, gscale);
    return (int)hipGetLastError();
}

// ---- internal interface for the Krylov driver (precond_api.hpp) ----------
#include "precond_api.hpp"

namespace bs {

static_assert(kUNetPartWg == kPartWg, "partial counts must agree");

size_t unet_ws_bytes(int nx, int ny, int nz) { return carve_unet(nullptr, nx, ny, nz, 0).bytes; }

int unet_apply(hipStream_t st, void* ws, int nx, int ny, int nz, const float* params, const void* packed,
               const float* r, const uint8_t* code, const unsigned* n_fluid, const double* rr_part, unsigned G,
               float* z, const int* done) {
    UNetWs W = carve_unet(ws, nx, ny, nz, 0);
    return unet_forward(st, W, params, (const __bf16*)packed, r, code, 0.0, n_fluid, rr_part, G, z,
                        done, 0);
}

}  // namespace bs
