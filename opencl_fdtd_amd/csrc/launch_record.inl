// Launch recording (part of the fdtd_hip.hip translation unit, compiled only
// with -DFDTD_RECORD_LAUNCHES, host code only).
//
// In that build every hipLaunchKernelGGL of the library is replaced by a
// record of the launch: the kernel, its grid and workgroup size, the
// march decomposition (Geom) and the buffers it reads and writes.  Nothing
// runs on a GPU and no field memory is touched, so the host logic (plan
// creation, launch shapes, the step_n schedule and its buffer swaps) can be
// exercised on a machine without a GPU and under AddressSanitizer and
// UBSan.  The product library is never built with it.

#include <vector>

extern "C" {

// Buffer slots of a record (component 0 of a field, or its packed base).
enum {
    FDTD_REC_E_IN = 0, FDTD_REC_E_OUT, FDTD_REC_H_IN, FDTD_REC_H_OUT, FDTD_REC_PSIE_IN, FDTD_REC_PSIE_OUT,
    FDTD_REC_PSIH_IN, FDTD_REC_PSIH_OUT, FDTD_REC_F_IN, FDTD_REC_F_OUT, FDTD_REC_G_IN, FDTD_REC_G_OUT,
    FDTD_REC_S, FDTD_REC_SRC, FDTD_REC_PTR0, FDTD_REC_SLOTS = 16
};

struct fdtd_launch_record {
    char kernel[128];
    int32_t grid, block;
    int32_t has_geom;
    int32_t sx, sy, sz, plane, xchunk, nchunk, pblocks, tile;
    int32_t src_x, src_p, src_comp;
    int32_t tyo, tzo, tzb, tw;      // he_march / he2_march tiles (Geom)
    const void* buf[FDTD_REC_SLOTS];
};

}  // extern "C"

namespace {

std::vector<fdtd_launch_record> g_records;
std::mutex g_records_mu;

template <class T>
void rec_geom(fdtd_launch_record& r, const Geom<T>& g) {
    r.has_geom = 1;
    r.sx = g.sx; r.sy = g.sy; r.sz = g.sz; r.plane = g.plane;
    r.xchunk = g.xchunk; r.nchunk = g.nchunk; r.pblocks = g.pblocks; r.tile = g.tile;
    r.tyo = g.tyo; r.tzo = g.tzo; r.tzb = g.tzb; r.tw = g.tw;
}

template <class T>
void rec_fill(fdtd_launch_record& r, const EArgs<T>& a) {
    rec_geom(r, a.g);
    r.buf[FDTD_REC_E_IN] = r.buf[FDTD_REC_E_OUT] = a.s[0].E[0];
    r.buf[FDTD_REC_H_IN] = a.s[0].H[0];
    r.buf[FDTD_REC_PSIE_IN] = r.buf[FDTD_REC_PSIE_OUT] = a.s[0].psi;
    r.buf[FDTD_REC_F_IN] = r.buf[FDTD_REC_F_OUT] = a.s[1].E[0];
    r.buf[FDTD_REC_G_IN] = a.s[1].H[0];
    r.buf[FDTD_REC_SRC] = a.src_amp;
    r.buf[FDTD_REC_PTR0] = a.ycorr[0];      // pending shell correction (null: none)
    r.src_x = a.src_x; r.src_p = a.src_p; r.src_comp = a.src_amp ? a.src_comp : -1;
}

template <class T>
void rec_fill(fdtd_launch_record& r, const HArgs<T>& a) {
    rec_geom(r, a.g);
    r.buf[FDTD_REC_E_IN] = a.s[0].E[0];
    r.buf[FDTD_REC_H_IN] = a.s[0].Hin[0];
    r.buf[FDTD_REC_H_OUT] = a.s[0].Hout[0];
    r.buf[FDTD_REC_PSIH_IN] = a.s[0].psi_in;
    r.buf[FDTD_REC_PSIH_OUT] = a.s[0].psi_out;
    r.buf[FDTD_REC_F_IN] = a.s[1].E[0];
    r.buf[FDTD_REC_G_IN] = a.s[1].Hin[0];
    r.buf[FDTD_REC_G_OUT] = a.s[1].Hout[0];
    r.buf[FDTD_REC_S] = a.S[0];
    r.src_comp = -1;
}

template <class T>
void rec_fill(fdtd_launch_record& r, const HEArgs<T>& a) {
    rec_geom(r, a.g);
    r.buf[FDTD_REC_E_IN] = a.hs.E.b;
    r.buf[FDTD_REC_E_OUT] = a.es.E.b;
    r.buf[FDTD_REC_H_IN] = a.hs.Hin.b;
    r.buf[FDTD_REC_H_OUT] = a.hs.Hout.b;
    r.buf[FDTD_REC_PSIE_IN] = r.buf[FDTD_REC_PSIE_OUT] = a.es.psi;
    r.buf[FDTD_REC_PSIH_IN] = a.hs.psi_in;
    r.buf[FDTD_REC_PSIH_OUT] = a.hs.psi_out;
    r.buf[FDTD_REC_S] = a.S.b;
    r.buf[FDTD_REC_SRC] = a.src_amp;
    r.buf[FDTD_REC_PTR0] = a.ycorr.b;
    r.src_x = a.src_x; r.src_p = a.src_p; r.src_comp = a.src_amp ? a.src_comp : -1;
}

template <class T>
void rec_fill(fdtd_launch_record& r, const HE2Args<T>& a) {
    rec_geom(r, a.g);
    r.buf[FDTD_REC_E_IN] = a.E.b;
    r.buf[FDTD_REC_E_OUT] = a.E_out.b;
    r.buf[FDTD_REC_H_IN] = a.H.b;
    r.buf[FDTD_REC_H_OUT] = a.H_out.b;
    r.buf[FDTD_REC_PSIE_IN] = a.psi_e_in;
    r.buf[FDTD_REC_PSIE_OUT] = a.psi_e_out;
    r.buf[FDTD_REC_PSIH_IN] = a.psi_h_in;
    r.buf[FDTD_REC_PSIH_OUT] = a.psi_h_out;
    r.buf[FDTD_REC_S] = a.S.b;
    r.buf[FDTD_REC_SRC] = a.src_amp;
    r.buf[FDTD_REC_PTR0] = a.ycorr.b;
    r.src_x = a.src_x; r.src_p = a.src_p; r.src_comp = a.src_amp ? a.src_comp : -1;
}

template <class T>
void rec_fill(fdtd_launch_record& r, const HEBArgs<T>& a) {
    rec_geom(r, a.g);
    r.buf[FDTD_REC_E_IN] = a.hs[0].E.b;
    r.buf[FDTD_REC_E_OUT] = a.es[0].E.b;
    r.buf[FDTD_REC_H_IN] = a.hs[0].Hin.b;
    r.buf[FDTD_REC_H_OUT] = a.hs[0].Hout.b;
    r.buf[FDTD_REC_PSIE_IN] = r.buf[FDTD_REC_PSIE_OUT] = a.es[0].psi;
    r.buf[FDTD_REC_PSIH_IN] = a.hs[0].psi_in;
    r.buf[FDTD_REC_PSIH_OUT] = a.hs[0].psi_out;
    r.buf[FDTD_REC_F_IN] = a.hs[1].E.b;
    r.buf[FDTD_REC_F_OUT] = a.es[1].E.b;
    r.buf[FDTD_REC_G_IN] = a.hs[1].Hin.b;
    r.buf[FDTD_REC_G_OUT] = a.hs[1].Hout.b;
    r.buf[FDTD_REC_S] = a.S.b;
    r.buf[FDTD_REC_SRC] = a.src_amp;
    r.src_x = a.src_x; r.src_p = a.src_p; r.src_comp = a.src_amp ? a.src_comp : -1;
}

template <class A>
void rec_fill(fdtd_launch_record& r, const A&) { r.src_comp = -1; }

// kernels with plain arguments (plane_sum, the reductions, add_point, ...):
// the first argument when it is a pointer
template <class T>
void rec_fill(fdtd_launch_record& r, T* const& p) {
    r.src_comp = -1;
    r.buf[FDTD_REC_PTR0] = p;
}

template <class A, class... Rest>
void fdtd_record_launch(const char* kernel, dim3 grid, dim3 block, const A& a, const Rest&...) {
    fdtd_launch_record r;
    memset(&r, 0, sizeof(r));
    snprintf(r.kernel, sizeof(r.kernel), "%s", kernel);
    r.grid = (int32_t)(grid.x * grid.y * grid.z);
    r.block = (int32_t)(block.x * block.y * block.z);
    rec_fill(r, a);
    std::lock_guard<std::mutex> lock(g_records_mu);
    g_records.push_back(r);
}

// A device-to-device copy: kernel "copy", source in E_IN, destination in
// E_OUT, the byte count in grid.
void fdtd_record_copy(void* dst, const void* src, size_t bytes) {
    fdtd_launch_record r;
    memset(&r, 0, sizeof(r));
    snprintf(r.kernel, sizeof(r.kernel), "copy");
    r.grid = (int32_t)std::min(bytes, (size_t)INT32_MAX);
    r.src_comp = -1;
    r.buf[FDTD_REC_E_IN] = src;
    r.buf[FDTD_REC_E_OUT] = dst;
    std::lock_guard<std::mutex> lock(g_records_mu);
    g_records.push_back(r);
}

}  // namespace

extern "C" {

int64_t fdtd_record_count(void) {
    std::lock_guard<std::mutex> lock(g_records_mu);
    return (int64_t)g_records.size();
}

int fdtd_record_get(int64_t i, fdtd_launch_record* out) {
    std::lock_guard<std::mutex> lock(g_records_mu);
    if (!out || i < 0 || i >= (int64_t)g_records.size()) return 1;
    *out = g_records[(size_t)i];
    return 0;
}

void fdtd_record_clear(void) {
    std::lock_guard<std::mutex> lock(g_records_mu);
    g_records.clear();
}

}  // extern "C"

#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, G, B, SHMEM, STREAM, ...) fdtd_record_launch(#K, G, B, __VA_ARGS__)
