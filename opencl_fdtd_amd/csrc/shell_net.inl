// The learned absorbing shell (part of the fdtd_hip.hip translation unit).
//
// A thin shell of T cells on every face runs the CPML-T update; after each E
// update (and source) a small convolutional net adds a correction to E on the
// shell cells:
//   x = (Ex, Ey, Ez, Hx, Hy, Hz)                         6 input channels
//   h_k(j) = sum_{c,t} W1[k][c][t] x_c(j + t)            16 hidden channels
//   y_a(i) = sum_{k,t} W2[a][k][t] h_k(i + t)            3 outputs
//   E_a(i) += y_a(i)  for i in the shell
// t runs over the 27 offsets of a 3x3x3 stencil (periodic wrap, as the step).
// There is no activation: the correction must be linear in the fields, as
// Maxwell's equations are.  Weights and activations are BF16 (inputs rounded
// to BF16), products are exact in FP32 and accumulated in FP32.
//
// The shell is split into 6 boxes, one per face (x faces own the edges and
// corners they touch, then y faces), each with its own weights.  Hidden values
// live per box on the box grown by one cell, in a caller-provided FP32
// scratch [cell][16].

namespace {

constexpr int kHid = 16;
constexpr int kIn = 6;
constexpr int kOut = 3;
constexpr int kTaps = 27;
constexpr int kW1 = kHid * kIn * kTaps;    // per box
constexpr int kW2 = kOut * kHid * kTaps;   // per box

// Along an axis the box spans completely (full), its grown regions are the
// same periodic range (local coordinate = global coordinate, wrapping);
// along the others they grow by 1 (hidden) or 2 (input adjoint) cells on
// each side.
struct ShellBox {
    int o0[3], o1[3];    // output region [o0, o1) per axis
    int full[3];         // 1: o0 = 0 and o1 = the grid size
    int s[3];            // grid size per axis
    int e[3];            // hidden (grown) extents
    int hoff;            // offset of the box's hidden block, in cells
    int hcells;          // e[0]*e[1]*e[2]
};

// global coordinate of hidden-local l, and hidden-local of global g (-1 if outside)
__device__ __forceinline__ int hid_global(const ShellBox& B, int q, int l) {
    return B.full[q] ? l : B.o0[q] - 1 + l;
}
__device__ __forceinline__ int hid_local(const ShellBox& B, int q, int g) {
    if (B.full[q]) return g < 0 ? g + B.s[q] : (g >= B.s[q] ? g - B.s[q] : g);
    const int l = g - (B.o0[q] - 1);
    return (l < 0 || l >= B.e[q]) ? -1 : l;
}
// output-local of global g (-1 if outside)
__device__ __forceinline__ int out_local(const ShellBox& B, int q, int g) {
    if (B.full[q]) return g < 0 ? g + B.s[q] : (g >= B.s[q] ? g - B.s[q] : g);
    const int l = g - B.o0[q];
    return (l < 0 || l >= B.o1[q] - B.o0[q]) ? -1 : l;
}
// extent and origin of the input-adjoint region (grown by 2)
__device__ __forceinline__ int in_extent(const ShellBox& B, int q) {
    return B.full[q] ? B.s[q] : B.o1[q] - B.o0[q] + 4;
}
__device__ __forceinline__ int in_global(const ShellBox& B, int q, int l) {
    return B.full[q] ? l : B.o0[q] - 2 + l;
}

template <class T>
struct ShellArgs {
    int sx, sy, sz, nbox;
    ShellBox box[6];
    const uint16_t* w;   // BF16: W1 for all boxes, then W2 for all boxes
    const T* E;          // input E (the correction reads E before it is applied)
    const T* H;
    T* Eout;             // E to correct (may equal E: layer 2 reads only hidden)
    float* hid;          // hidden scratch
    float* hbar;         // hidden adjoint scratch
    const T* Ebar;       // adjoint of corrected E (backward)
    T* Ebar_out;         // adjoint of E before the correction (accumulated)
    T* Hbar_out;         // adjoint of H (accumulated)
    float* wgrad_part;   // per-workgroup weight-gradient partials
    int part_box;        // box handled by a per-box launch
    int nparts;
};

__device__ __forceinline__ float bf16_to_f(uint16_t v) {
    return __uint_as_float((uint32_t)v << 16);
}
// round-to-nearest-even to BF16, returned as the FP32 value it represents
__device__ __forceinline__ float bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;   // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ int wrapi(int c, int s) { return c < 0 ? c + s : (c >= s ? c - s : c); }

template <class T>
__device__ __forceinline__ size_t gidx(const ShellArgs<T>& a, int x, int y, int z) {
    return ((size_t)wrapi(x, a.sx) * a.sy + wrapi(y, a.sy)) * a.sz + wrapi(z, a.sz);
}

// global linear cell of the grown box -> (box, local coords)
template <class T>
__device__ __forceinline__ bool locate(const ShellArgs<T>& a, int cell, int* b, int* lx, int* ly, int* lz) {
    for (int k = 0; k < a.nbox; ++k) {
        const ShellBox& B = a.box[k];
        if (cell >= B.hoff && cell < B.hoff + B.hcells) {
            const int l = cell - B.hoff;
            *b = k;
            *lz = l % B.e[2];
            *ly = (l / B.e[2]) % B.e[1];
            *lx = l / (B.e[2] * B.e[1]);
            return true;
        }
    }
    return false;
}

// layer 1 over every grown-box cell
template <class T>
__global__ __launch_bounds__(256) void shell_l1(ShellArgs<T> a, int total) {
    const int cell = blockIdx.x * 256 + threadIdx.x;
    int b, lx, ly, lz;
    if (cell >= total || !locate(a, cell, &b, &lx, &ly, &lz)) return;
    const ShellBox& B = a.box[b];
    const int x = hid_global(B, 0, lx), y = hid_global(B, 1, ly), z = hid_global(B, 2, lz);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w1 = a.w + (size_t)b * kW1;
    float acc[kHid];
#pragma unroll
    for (int k = 0; k < kHid; ++k) acc[k] = 0.f;
    for (int t = 0; t < kTaps; ++t) {
        const size_t gi = gidx(a, x + t / 9 - 1, y + (t / 3) % 3 - 1, z + t % 3 - 1);
        float xv[kIn];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xv[c] = bf16_round((float)a.E[c * N + gi]);
            xv[3 + c] = bf16_round((float)a.H[c * N + gi]);
        }
#pragma unroll
        for (int k = 0; k < kHid; ++k)
#pragma unroll
            for (int c = 0; c < kIn; ++c) acc[k] += bf16_to_f(w1[(k * kIn + c) * kTaps + t]) * xv[c];
    }
    float* h = a.hid + (size_t)cell * kHid;
#pragma unroll
    for (int k = 0; k < kHid; ++k) h[k] = bf16_round(acc[k]);
}

// layer 2 over the output cells, adding the correction to Eout
template <class T>
__global__ __launch_bounds__(256) void shell_l2(ShellArgs<T> a, int total_out) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= total_out) return;
    int b = 0;
    for (; b < a.nbox; ++b) {
        const ShellBox& B = a.box[b];
        const int n = (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
        if (cell < n) break;
        cell -= n;
    }
    if (b >= a.nbox) return;
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int lz = cell % nz, ly = (cell / nz) % ny, lx = cell / (nz * ny);
    const uint16_t* w2 = a.w + (size_t)a.nbox * kW1 + (size_t)b * kW2;
    float y[kOut] = {0.f, 0.f, 0.f};
    for (int t = 0; t < kTaps; ++t) {
        const int hx = hid_local(B, 0, B.o0[0] + lx + t / 9 - 1);
        const int hy = hid_local(B, 1, B.o0[1] + ly + (t / 3) % 3 - 1);
        const int hz = hid_local(B, 2, B.o0[2] + lz + t % 3 - 1);
        const float* h = a.hid + ((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid;
#pragma unroll
        for (int k = 0; k < kHid; ++k) {
            const float hv = h[k];
#pragma unroll
            for (int o = 0; o < kOut; ++o) y[o] += bf16_to_f(w2[(o * kHid + k) * kTaps + t]) * hv;
        }
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz);
#pragma unroll
    for (int o = 0; o < kOut; ++o) a.Eout[o * N + gi] += (T)y[o];
}

// ---- backward (training): adjoint of E_out = E + mask * y(E, H) ----------

// hbar_k(j) = sum_{o,t} W2[o][k][t] ybar_o(j - t), ybar = Ebar on the box's
// output cells (0 elsewhere); j over the grown box
template <class T>
__global__ __launch_bounds__(256) void shell_l2_t(ShellArgs<T> a, int total) {
    const int cell = blockIdx.x * 256 + threadIdx.x;
    int b, lx, ly, lz;
    if (cell >= total || !locate(a, cell, &b, &lx, &ly, &lz)) return;
    const ShellBox& B = a.box[b];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w2 = a.w + (size_t)a.nbox * kW1 + (size_t)b * kW2;
    float acc[kHid];
#pragma unroll
    for (int k = 0; k < kHid; ++k) acc[k] = 0.f;
    const int gx = hid_global(B, 0, lx), gy = hid_global(B, 1, ly), gz = hid_global(B, 2, lz);
    for (int t = 0; t < kTaps; ++t) {
        // output cell i = j - t
        const int ox = out_local(B, 0, gx - (t / 9 - 1));
        const int oy = out_local(B, 1, gy - ((t / 3) % 3 - 1));
        const int oz = out_local(B, 2, gz - (t % 3 - 1));
        if (ox < 0 || oy < 0 || oz < 0) continue;
        const size_t gi = gidx(a, B.o0[0] + ox, B.o0[1] + oy, B.o0[2] + oz);
        float yb[kOut];
#pragma unroll
        for (int o = 0; o < kOut; ++o) yb[o] = (float)a.Ebar[o * N + gi];
#pragma unroll
        for (int k = 0; k < kHid; ++k)
#pragma unroll
            for (int o = 0; o < kOut; ++o) acc[k] += bf16_to_f(w2[(o * kHid + k) * kTaps + t]) * yb[o];
    }
    float* hb = a.hbar + (size_t)cell * kHid;
#pragma unroll
    for (int k = 0; k < kHid; ++k) hb[k] = acc[k];
}

// xbar_c(i) = sum_{k,t} W1[k][c][t] hbar_k(i - t) for one box (part_box);
// i over the box grown by two cells; accumulated into Ebar_out / Hbar_out.
template <class T>
__global__ __launch_bounds__(256) void shell_l1_t(ShellArgs<T> a) {
    const ShellBox& B = a.box[a.part_box];
    const int g0 = in_extent(B, 0), g1 = in_extent(B, 1), g2 = in_extent(B, 2);
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= g0 * g1 * g2) return;
    const int lz = cell % g2, ly = (cell / g2) % g1, lx = cell / (g2 * g1);   // grown-by-2 coords
    const int gx = in_global(B, 0, lx), gy = in_global(B, 1, ly), gz = in_global(B, 2, lz);
    const uint16_t* w1 = a.w + (size_t)a.part_box * kW1;
    float acc[kIn];
#pragma unroll
    for (int c = 0; c < kIn; ++c) acc[c] = 0.f;
    for (int t = 0; t < kTaps; ++t) {
        // hidden cell j = i - t
        const int hx = hid_local(B, 0, gx - (t / 9 - 1));
        const int hy = hid_local(B, 1, gy - ((t / 3) % 3 - 1));
        const int hz = hid_local(B, 2, gz - (t % 3 - 1));
        if (hx < 0 || hy < 0 || hz < 0) continue;
        const float* hb = a.hbar + ((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid;
#pragma unroll
        for (int k = 0; k < kHid; ++k) {
            const float v = hb[k];
#pragma unroll
            for (int c = 0; c < kIn; ++c) acc[c] += bf16_to_f(w1[(k * kIn + c) * kTaps + t]) * v;
        }
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, gx, gy, gz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.Ebar_out[c * N + gi] += (T)acc[c];
        a.Hbar_out[c * N + gi] += (T)acc[3 + c];
    }
}

// Weight-gradient partials of one box (part_box), GEMM-shaped:
//   W1bar[k][(c,t)] = sum_j hbar_k(j) x_c(j + t)   over the grown box
//   W2bar[o][(k,t)] = sum_i ybar_o(i) h_k(i + t)   over the output box
// Workgroup g takes a contiguous slice of the cells; chunks of the slice are
// staged in LDS (adjoint rows and im2col rows) and each thread accumulates a
// fixed subset of the weights.  partials[g][w], summed in a fixed order later.
constexpr int kChunk1 = 32;
constexpr int kChunk2 = 16;

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad1(ShellArgs<T> a) {
    __shared__ float sh[kChunk1][kHid];
    __shared__ float sx[kChunk1][kIn * kTaps];
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int per = (B.hcells + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, B.hcells);
    constexpr int NW = kHid * kIn * kTaps;   // 2592
    constexpr int PER = (NW + 255) / 256;
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.f;
    for (int base = c0; base < c1; base += kChunk1) {
        const int n = min(kChunk1, c1 - base);
        for (int e = threadIdx.x; e < kChunk1 * kHid; e += 256) {
            const int l = e / kHid, k = e % kHid;
            sh[l][k] = l < n ? a.hbar[((size_t)B.hoff + base + l) * kHid + k] : 0.f;
        }
        for (int e = threadIdx.x; e < kChunk1 * kIn * kTaps; e += 256) {
            const int l = e / (kIn * kTaps), ct = e % (kIn * kTaps), c = ct / kTaps, t = ct % kTaps;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % B.e[2], ly = (cl / B.e[2]) % B.e[1], lx = cl / (B.e[2] * B.e[1]);
                const size_t gi = gidx(a, hid_global(B, 0, lx) + t / 9 - 1, hid_global(B, 1, ly) + (t / 3) % 3 - 1,
                                       hid_global(B, 2, lz) + t % 3 - 1);
                v = bf16_round((float)(c < 3 ? a.E[c * N + gi] : a.H[(c - 3) * N + gi]));
            }
            sx[l][ct] = v;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int wi = threadIdx.x + 256 * m;
            if (wi < NW) {
                const int k = wi / (kIn * kTaps), ct = wi % (kIn * kTaps);
                float s_ = 0.f;
                for (int l = 0; l < kChunk1; ++l) s_ += sh[l][k] * sx[l][ct];
                acc[m] += s_;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int wi = threadIdx.x + 256 * m;
        if (wi < NW) a.wgrad_part[(size_t)blockIdx.x * NW + wi] = acc[m];
    }
}

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad2(ShellArgs<T> a) {
    __shared__ float sy[kChunk2][kOut];
    __shared__ float shc[kChunk2][kHid * kTaps];
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int nout = (B.o1[0] - B.o0[0]) * ny * nz;
    const int per = (nout + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, nout);
    constexpr int NW = kOut * kHid * kTaps;  // 1296
    constexpr int PER = (NW + 255) / 256;
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.f;
    for (int base = c0; base < c1; base += kChunk2) {
        const int n = min(kChunk2, c1 - base);
        for (int e = threadIdx.x; e < kChunk2 * kOut; e += 256) {
            const int l = e / kOut, o = e % kOut;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % nz, ly = (cl / nz) % ny, lx = cl / (nz * ny);
                v = (float)a.Ebar[o * N + gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz)];
            }
            sy[l][o] = v;
        }
        for (int e = threadIdx.x; e < kChunk2 * kHid * kTaps; e += 256) {
            const int l = e / (kHid * kTaps), kt = e % (kHid * kTaps), k = kt / kTaps, t = kt % kTaps;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % nz, ly = (cl / nz) % ny, lx = cl / (nz * ny);
                const int hx = hid_local(B, 0, B.o0[0] + lx + t / 9 - 1);
                const int hy = hid_local(B, 1, B.o0[1] + ly + (t / 3) % 3 - 1);
                const int hz = hid_local(B, 2, B.o0[2] + lz + t % 3 - 1);
                v = a.hid[((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid + k];
            }
            shc[l][kt] = v;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int wi = threadIdx.x + 256 * m;
            if (wi < NW) {
                const int o = wi / (kHid * kTaps), kt = wi % (kHid * kTaps);
                float s_ = 0.f;
                for (int l = 0; l < kChunk2; ++l) s_ += sy[l][o] * shc[l][kt];
                acc[m] += s_;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int wi = threadIdx.x + 256 * m;
        if (wi < NW) a.wgrad_part[(size_t)blockIdx.x * NW + wi] = acc[m];
    }
}

// wgrad[b][w] += sum_g partials[g][w], fixed order
__global__ __launch_bounds__(256) void shell_wgrad_sum(const float* part, int groups, int nw, float* out) {
    const int wi = blockIdx.x * 256 + threadIdx.x;
    if (wi >= nw) return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += (double)part[(size_t)g * nw + wi];
    out[wi] += (float)s;
}

// ---- inference with MFMA (v_mfma_f32_16x16x32_bf16) ----------------------
//
// One workgroup (4 waves) per tile of a box: T x 8 x 16 output cells (the
// thin axis first).  The input fields of the tile grown by 2 are staged in
// LDS as BF16 [cell][8] (6 channels + 2 zero pads), the hidden layer of the
// tile grown by 1 as BF16 [cell][16].  Both layers are GEMMs over (tap,
// channel):
//   layer 1: M = hidden cells, K = 28 taps x 8 channels (224; tap 27 and
//            channels 6, 7 are zero), N = 16 hidden channels;
//   layer 2: M = output cells, K = 28 taps x 16 hidden channels, N = 16
//            (3 outputs used).
// MFMA operands (16x16x32): lane l holds A row l%16 and B column l%16 at
// k = 8*(l/16) + [0..7]; D column l%16 at rows 4*(l/16) + [0..3].  With the
// K order above, a lane's 8 A values are one 16-byte LDS read.  The
// correction is written to ybuf[out cell][3] and added to E by shell_add,
// since neighbouring tiles read E around their own cells.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileA = 8, kTileB = 16;
constexpr int kMaxT = 4;
constexpr int kInMax = (kMaxT + 4) * (kTileA + 4) * (kTileB + 4);
constexpr int kHidMax = (kMaxT + 2) * (kTileA + 2) * (kTileB + 2);

struct ShellTiles {
    int te[6][3];      // tile extents per axis
    int nt[6][3];      // tiles per axis
    int tstart[7];     // prefix of tiles per box
    int ostart[6];     // prefix of output cells per box
};

__device__ __forceinline__ uint16_t f2bf(float f) {
    return (uint16_t)(__float_as_uint(bf16_round(f)) >> 16);
}

template <class T>
__global__ __launch_bounds__(256) void shell_mfma(ShellArgs<T> a, ShellTiles tl, float* ybuf) {
    __shared__ uint4 s_in[kInMax];
    __shared__ uint4 s_hid[kHidMax * 2];
    int b = 0;
    while (b < a.nbox - 1 && (int)blockIdx.x >= tl.tstart[b + 1]) ++b;
    const ShellBox& B = a.box[b];
    int tix = blockIdx.x - tl.tstart[b];
    int tcoord[3];
    tcoord[2] = tix % tl.nt[b][2];
    tcoord[1] = (tix / tl.nt[b][2]) % tl.nt[b][1];
    tcoord[0] = tix / (tl.nt[b][2] * tl.nt[b][1]);
    int o[3], te[3], ie[3], he[3];
    for (int q = 0; q < 3; ++q) {
        te[q] = tl.te[b][q];
        o[q] = B.o0[q] + tcoord[q] * te[q];
        ie[q] = te[q] + 4;
        he[q] = te[q] + 2;
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int nin = ie[0] * ie[1] * ie[2];
    const int nhid = he[0] * he[1] * he[2];
    const int nout = te[0] * te[1] * te[2];
    // 1. input tile -> LDS (BF16, 8 per cell)
    for (int c = threadIdx.x; c < nin; c += 256) {
        const int lz = c % ie[2], ly = (c / ie[2]) % ie[1], lx = c / (ie[2] * ie[1]);
        const size_t gi = gidx(a, o[0] - 2 + lx, o[1] - 2 + ly, o[2] - 2 + lz);
        uint32_t w[4];
        const uint16_t v0 = f2bf((float)a.E[gi]), v1 = f2bf((float)a.E[N + gi]), v2 = f2bf((float)a.E[2 * N + gi]);
        const uint16_t v3 = f2bf((float)a.H[gi]), v4 = f2bf((float)a.H[N + gi]), v5 = f2bf((float)a.H[2 * N + gi]);
        w[0] = (uint32_t)v0 | ((uint32_t)v1 << 16);
        w[1] = (uint32_t)v2 | ((uint32_t)v3 << 16);
        w[2] = (uint32_t)v4 | ((uint32_t)v5 << 16);
        w[3] = 0;
        s_in[c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    // weight fragments of this lane
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, n = lane & 15;
    const uint16_t* w1 = a.w + (size_t)b * kW1;
    const uint16_t* w2 = a.w + (size_t)a.nbox * kW1 + (size_t)b * kW2;
    bf16x8 B1[7], B2[14];
    for (int j = 0; j < 7; ++j) {
        const int tap = 4 * j + g;
        for (int i = 0; i < 8; ++i) {
            const uint16_t v = (tap < kTaps && i < kIn) ? w1[(n * kIn + i) * kTaps + tap] : (uint16_t)0;
            B1[j][i] = __builtin_bit_cast(__bf16, v);
        }
    }
    for (int j = 0; j < 14; ++j) {
        const int tap = 2 * j + (g >> 1);
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * (g & 1) + i;
            const uint16_t v = (tap < kTaps && n < kOut) ? w2[(n * kHid + k) * kTaps + tap] : (uint16_t)0;
            B2[j][i] = __builtin_bit_cast(__bf16, v);
        }
    }
    __syncthreads();
    // 2. layer 1: hidden = conv1(input)
    const int mt1 = (nhid + 15) / 16;
    uint16_t* hid16 = reinterpret_cast<uint16_t*>(s_hid);
    for (int mt = wave; mt < mt1; mt += 4) {
        const int row = min(mt * 16 + n, nhid - 1);
        const int hz = row % he[2], hy = (row / he[2]) % he[1], hx = row / (he[2] * he[1]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int tap = min(4 * j + g, kTaps - 1);     // tap 27 has zero weights
            const int ix = hx + 1 + tap / 9 - 1, iy = hy + 1 + (tap / 3) % 3 - 1, iz = hz + 1 + tap % 3 - 1;
            const uint4 v = s_in[(ix * ie[1] + iy) * ie[2] + iz];
            const bf16x8 av = __builtin_bit_cast(bf16x8, v);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, B1[j], acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = mt * 16 + 4 * g + i;
            if (r < nhid) hid16[r * kHid + n] = f2bf(acc[i]);
        }
    }
    __syncthreads();
    // 3. layer 2: y = conv2(hidden) on the tile's output cells
    const int mt2 = (nout + 15) / 16;
    for (int mt = wave; mt < mt2; mt += 4) {
        const int row = min(mt * 16 + n, nout - 1);
        const int oz = row % te[2], oy = (row / te[2]) % te[1], ox = row / (te[2] * te[1]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 14; ++j) {
            const int tap = min(2 * j + (g >> 1), kTaps - 1);
            const int hx = ox + 1 + tap / 9 - 1, hy = oy + 1 + (tap / 3) % 3 - 1, hz = oz + 1 + tap % 3 - 1;
            const uint4 v = s_hid[((hx * he[1] + hy) * he[2] + hz) * 2 + (g & 1)];
            const bf16x8 av = __builtin_bit_cast(bf16x8, v);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, B2[j], acc, 0, 0, 0);
        }
        if (n < kOut) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = mt * 16 + 4 * g + i;
                if (r >= nout) continue;
                const int rz = r % te[2], ry = (r / te[2]) % te[1], rx = r / (te[2] * te[1]);
                const int gx = o[0] + rx, gy = o[1] + ry, gz = o[2] + rz;
                if (gx >= B.o1[0] || gy >= B.o1[1] || gz >= B.o1[2]) continue;
                const int bx = gx - B.o0[0], by = gy - B.o0[1], bz = gz - B.o0[2];
                const int oi = tl.ostart[b] + (bx * (B.o1[1] - B.o0[1]) + by) * (B.o1[2] - B.o0[2]) + bz;
                ybuf[(size_t)oi * kOut + n] = acc[i];
            }
        }
    }
}

// E += ybuf on the shell cells
template <class T>
__global__ __launch_bounds__(256) void shell_add(ShellArgs<T> a, const float* ybuf, int total_out) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= total_out) return;
    const int oi = cell;
    int b = 0;
    for (; b < a.nbox; ++b) {
        const ShellBox& B = a.box[b];
        const int n = (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
        if (cell < n) break;
        cell -= n;
    }
    if (b >= a.nbox) return;
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int lz = cell % nz, ly = (cell / nz) % ny, lx = cell / (nz * ny);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz);
#pragma unroll
    for (int o = 0; o < kOut; ++o) a.Eout[o * N + gi] += (T)ybuf[(size_t)oi * kOut + o];
}

struct ShellPlan {
    int nbox;
    ShellBox box[6];
    int hcells, out_cells;
    ShellTiles tiles;
    int ntiles;
};

ShellPlan make_shell(const Plan& P, int T) {
    ShellPlan S;
    const int s[3] = {P.sx, P.sy, P.sz};
    int nb = 0;
    for (int r = 0; r < 3; ++r) {
        for (int side = 0; side < 2; ++side) {
            ShellBox& B = S.box[nb++];
            for (int q = 0; q < 3; ++q) {
                if (q < r) { B.o0[q] = T; B.o1[q] = s[q] - T; }
                else if (q > r) { B.o0[q] = 0; B.o1[q] = s[q]; }
                else { B.o0[q] = side == 0 ? 0 : s[q] - T; B.o1[q] = side == 0 ? T : s[q]; }
            }
        }
    }
    S.nbox = nb;
    int off = 0, outc = 0;
    for (int b = 0; b < nb; ++b) {
        ShellBox& B = S.box[b];
        for (int q = 0; q < 3; ++q) {
            B.s[q] = s[q];
            B.full[q] = (B.o0[q] == 0 && B.o1[q] == s[q]) ? 1 : 0;
            B.e[q] = B.full[q] ? s[q] : B.o1[q] - B.o0[q] + 2;
        }
        B.hcells = B.e[0] * B.e[1] * B.e[2];
        B.hoff = off;
        off += B.hcells;
        outc += (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
    }
    S.hcells = off;
    S.out_cells = outc;
    // MFMA tiles: the box's thin axis takes T, the other two 8 and 16
    int ts = 0, os = 0;
    for (int b = 0; b < nb; ++b) {
        const int r = b / 2;
        int k = 0;
        for (int q = 0; q < 3; ++q) {
            const int te = (q == r) ? T : (k++ == 0 ? kTileA : kTileB);
            const int ext = S.box[b].o1[q] - S.box[b].o0[q];
            S.tiles.te[b][q] = te;
            S.tiles.nt[b][q] = (ext + te - 1) / te;
        }
        S.tiles.tstart[b] = ts;
        S.tiles.ostart[b] = os;
        ts += S.tiles.nt[b][0] * S.tiles.nt[b][1] * S.tiles.nt[b][2];
        os += (S.box[b].o1[0] - S.box[b].o0[0]) * (S.box[b].o1[1] - S.box[b].o0[1]) * (S.box[b].o1[2] - S.box[b].o0[2]);
    }
    S.tiles.tstart[nb] = ts;
    S.ntiles = ts;
    return S;
}

template <class T>
void fill_shell(const Plan& P, const ShellPlan& S, ShellArgs<T>& a) {
    memset(&a, 0, sizeof(a));
    a.sx = P.sx; a.sy = P.sy; a.sz = P.sz;
    a.nbox = S.nbox;
    for (int b = 0; b < S.nbox; ++b) a.box[b] = S.box[b];
}

// T >= 2 keeps every grown region free of self-overlap through the wrap;
// T <= 4 bounds the MFMA tiles' LDS.
int check_shell(const Plan& P, int T) {
    if (T < 2 || T > kMaxT) return fail(3, "shell thickness %d not in [2, %d]", T, kMaxT);
    if (2 * T + 4 > P.sx || 2 * T + 4 > P.sy || 2 * T + 4 > P.sz)
        return fail(3, "shell thickness %d too large for the grid", T);
    return 0;
}

}  // namespace

extern "C" {

int64_t fdtd_shell_hidden_elems(const void* plan, int32_t thickness) {
    if (!plan) return -1;
    const Plan& P = *(const Plan*)plan;
    if (check_shell(P, thickness)) return -1;
    return (int64_t)make_shell(P, thickness).hcells * kHid;
}

int64_t fdtd_shell_weight_elems(void) { return 6 * (int64_t)(kW1 + kW2); }

int fdtd_shell_apply(const void* plan, int32_t thickness, const uint16_t* w, const void* E, const void* H,
                     float* hidden, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !E || !H || !hidden) return fail(2, "w, E, H and hidden are required");
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    const int g2 = (S.out_cells + 255) / 256;
    // hidden scratch holds the correction (out_cells x 3 <= hcells x 16)
    if (P.d.dtype == FDTD_F32) {
        ShellArgs<float> a;
        fill_shell<float>(P, S, a);
        a.w = w; a.E = (const float*)E; a.H = (const float*)H; a.Eout = (float*)E;
        hipLaunchKernelGGL(shell_mfma<float>, dim3(S.ntiles), dim3(256), 0, st, a, S.tiles, hidden);
        hipLaunchKernelGGL(shell_add<float>, dim3(g2), dim3(256), 0, st, a, (const float*)hidden, S.out_cells);
    } else {
        ShellArgs<double> a;
        fill_shell<double>(P, S, a);
        a.w = w; a.E = (const double*)E; a.H = (const double*)H; a.Eout = (double*)E;
        hipLaunchKernelGGL(shell_mfma<double>, dim3(S.ntiles), dim3(256), 0, st, a, S.tiles, hidden);
        hipLaunchKernelGGL(shell_add<double>, dim3(g2), dim3(256), 0, st, a, (const float*)hidden, S.out_cells);
    }
    return check_launch("fdtd_shell_apply");
}

int64_t fdtd_shell_partials_elems(void) { return (int64_t)64 * kW1; }

int fdtd_shell_backward(const void* plan, int32_t thickness, const uint16_t* w, const void* E, const void* H,
                        float* hidden, float* hbar, const void* Ebar, void* Ebar_out, void* Hbar_out,
                        float* partials, float* wgrad, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !E || !H || !hidden || !hbar || !Ebar || !Ebar_out || !Hbar_out || !partials || !wgrad)
        return fail(2, "all buffers are required");
    if (Ebar == Ebar_out) return fail(2, "Ebar_out must differ from Ebar");
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    const int groups = 64;
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        ShellArgs<T> a;
        fill_shell<T>(P, S, a);
        a.w = w; a.E = (const T*)E; a.H = (const T*)H; a.hid = hidden; a.hbar = hbar;
        a.Ebar = (const T*)Ebar; a.Ebar_out = (T*)Ebar_out; a.Hbar_out = (T*)Hbar_out;
        a.wgrad_part = partials;
        const int g1 = (S.hcells + 255) / 256;
        // forward hidden (the weight gradient of layer 2 needs it)
        hipLaunchKernelGGL(shell_l1<T>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        hipLaunchKernelGGL(shell_l2_t<T>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        for (int b = 0; b < S.nbox; ++b) {
            a.part_box = b;
            const ShellBox& B = S.box[b];
            const int n2 = (B.full[0] ? B.s[0] : B.o1[0] - B.o0[0] + 4) *
                           (B.full[1] ? B.s[1] : B.o1[1] - B.o0[1] + 4) *
                           (B.full[2] ? B.s[2] : B.o1[2] - B.o0[2] + 4);
            hipLaunchKernelGGL(shell_l1_t<T>, dim3((n2 + 255) / 256), dim3(256), 0, st, a);
            // W1 gradient of box b
            hipLaunchKernelGGL(shell_wgrad1<T>, dim3(groups), dim3(256), 0, st, a);
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kW1 + 255) / 256), dim3(256), 0, st, (const float*)partials,
                               groups, kW1, wgrad + (size_t)b * kW1);
            // W2 gradient of box b
            hipLaunchKernelGGL(shell_wgrad2<T>, dim3(groups), dim3(256), 0, st, a);
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kW2 + 255) / 256), dim3(256), 0, st, (const float*)partials,
                               groups, kW2, wgrad + (size_t)S.nbox * kW1 + (size_t)b * kW2);
        }
        return 0;
    };
    if (P.d.dtype == FDTD_F32) run(0.f);
    else run(0.0);
    return check_launch("fdtd_shell_backward");
}

}  // extern "C"
