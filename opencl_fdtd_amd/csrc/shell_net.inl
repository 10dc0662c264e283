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

struct ShellPlan {
    int nbox;
    ShellBox box[6];
    int hcells, out_cells;
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
    return S;
}

template <class T>
void fill_shell(const Plan& P, const ShellPlan& S, ShellArgs<T>& a) {
    memset(&a, 0, sizeof(a));
    a.sx = P.sx; a.sy = P.sy; a.sz = P.sz;
    a.nbox = S.nbox;
    for (int b = 0; b < S.nbox; ++b) a.box[b] = S.box[b];
}

// T >= 2 keeps every grown region free of self-overlap through the wrap.
int check_shell(const Plan& P, int T) {
    if (T < 2) return fail(3, "shell thickness %d < 2", T);
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
    const int g1 = (S.hcells + 255) / 256, g2 = (S.out_cells + 255) / 256;
    if (P.d.dtype == FDTD_F32) {
        ShellArgs<float> a;
        fill_shell<float>(P, S, a);
        a.w = w; a.E = (const float*)E; a.H = (const float*)H; a.Eout = (float*)E; a.hid = hidden;
        hipLaunchKernelGGL(shell_l1<float>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        hipLaunchKernelGGL(shell_l2<float>, dim3(g2), dim3(256), 0, st, a, S.out_cells);
    } else {
        ShellArgs<double> a;
        fill_shell<double>(P, S, a);
        a.w = w; a.E = (const double*)E; a.H = (const double*)H; a.Eout = (double*)E; a.hid = hidden;
        hipLaunchKernelGGL(shell_l1<double>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        hipLaunchKernelGGL(shell_l2<double>, dim3(g2), dim3(256), 0, st, a, S.out_cells);
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
