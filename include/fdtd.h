/*
 * fdtd.h -- C-ABI of the MI355X FDTD time-stepping library (libfdtd_hip.so).
 *
 * Plain C: device pointers, sizes and a stream handle; no torch types.  The
 * caller (the Python Simulation class, through ctypes) owns every device
 * buffer; the library allocates no device memory.
 *
 * Each entry point replaces one piece of the reference's device boundary
 * (opencl_fdtd, where every update is a pyopencl ElementwiseKernel built from
 * a Jinja2 template and bound with its argument list at
 * opencl_fdtd/simulation.py:321-326):
 *
 *   fdtd_plan_create   <- template rendering + ElementwiseKernel construction,
 *                         simulation.py:170-271 (shape baked in at
 *                         kernels/common.cl:16-19, PML list at :161-168,
 *                         coefficients at :274-319)
 *   fdtd_update_e      <- sim.update_E (+ sim.update_F with Bloch),
 *                         kernels/update_e.cl, bound at simulation.py:252,263
 *   fdtd_update_h      <- sim.update_H (+ sim.update_G), kernels/update_h.cl,
 *                         bound at simulation.py:253,264; with Poynting fused
 *                         it also does the work of kernels/update_s.cl
 *   fdtd_update_s      <- sim.update_S, kernels/update_s.cl, simulation.py:260-261
 *   fdtd_update_j      <- sim.update_J, kernels/update_j.cl, simulation.py:265-271
 *   fdtd_add_point     <- the caller's soft source  sim.E[ind] += f(t)
 *                         (fdtd.py:168-169), with f(t) read from device memory
 *   fdtd_step_n        <- n iterations of the caller loop fdtd.py:162-175
 *   fdtd_step_n_moved     (E, F, source, H, G, S) without host round trips
 *   fdtd_step_n_observed  (the same with a plane flux and field copies per step)
 *   fdtd_step_n_chain     (the same, the H half-step of the loop's last
 *                         iteration left for the next call, as sim.update_H
 *                         defers into the next sim.update_E)
 *   fdtd_update_he     <- sim.update_H + sim.update_S of one iteration of that
 *                         loop followed by sim.update_E and the soft source of
 *                         the next, in one pass over the grid
 *   fdtd_adjoint_step, fdtd_plane_sum, fdtd_energy, fdtd_reduce_partials
 *                      <- no reference counterpart (gradients and FP64
 *                         reductions over the same step; see below)
 *   fdtd_shell_*       <- no reference counterpart (the learned 4-cell
 *                         absorbing shell that stands in for a thick CPML;
 *                         see below)
 *
 * Field layout (kernels/common.cl:34-61, fdfd_tools.vec at simulation.py:347):
 * a vector field is 3 contiguous components of N = sx*sy*sz values each, and
 * inside a component index i = z + sz*(y + sy*x).  The domain wraps
 * periodically on every axis (kernels/common.cl:75-85).
 *
 * Every call returns 0 on success or a nonzero status; fdtd_last_error()
 * then describes the failure (thread-local).  Calls are asynchronous on the
 * given stream (a hipStream_t; NULL = the null stream).  A plan is immutable
 * after creation and may be shared by streams; buffers may not.
 */
#ifndef FDTD_MI355X_H
#define FDTD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_ABI_VERSION 7
#define FDTD_MAX_FACES 12

enum { FDTD_F32 = 0, FDTD_F64 = 1 };

/* Poynting modes (do_poynting in the reference, simulation.py:87, 227-235) */
enum {
    FDTD_POYNTING_OFF = 0,
    FDTD_POYNTING_FUSED = 1,     /* S computed inside fdtd_update_h; needs H != H_out */
    FDTD_POYNTING_TWO_PASS = 2   /* reference dataflow: H writes oS (6N), fdtd_update_s reads it */
};

/* Bloch modes (reference: simulation.py:157-159, 187-224) */
enum {
    FDTD_BLOCH_PHASE = 0,        /* neighbour across the seam multiplied by the phase factor */
    FDTD_BLOCH_DECOUPLED = 1     /* reference as shipped: F/G a second periodic system */
};

/* One CPML face, in list order (simulation.py:161-168, 274-319). */
typedef struct {
    int32_t axis;        /* 0 = x, 1 = y, 2 = z */
    int32_t polarity;    /* 0 = 'n' (low side), 1 = 'p' (high side) */
    int32_t thickness;   /* layers T */
    int32_t reserved;
} fdtd_face;

typedef struct {
    int64_t shape[3];        /* sx, sy, sz */
    int32_t dtype;           /* FDTD_F32 or FDTD_F64 */
    int32_t n_face;          /* <= FDTD_MAX_FACES */
    fdtd_face face[FDTD_MAX_FACES];
    double dt;               /* passed to the kernels as the field type */
    double inv_dx;           /* uniform 1/dx (computed in double, rounded to the field type) */
    int32_t nonuniform;      /* 1: per-index 1/dx arrays in fdtd_fields.inv_d */
    int32_t poynting;        /* FDTD_POYNTING_* */
    int32_t bloch_mask;      /* bit r set: Bloch boundary on axis r (needs F, G) */
    int32_t bloch_mode;      /* FDTD_BLOCH_* */
    double bloch_re[3];      /* phase factor real(exp(i phi)) per axis */
    double bloch_im[3];      /* phase factor imag(exp(i phi)) per axis */
} fdtd_desc;

/* Device buffers of one simulation.  Vector fields hold 3N values. */
typedef struct {
    void* E;          /* E field, updated in place by fdtd_update_e */
    void* H;          /* H field read by the H step (H^{n-1/2}) */
    void* H_out;      /* H field written by the H step; == H for in-place.
                         FDTD_POYNTING_FUSED needs H_out != H (ping-pong) */
    void* eps;        /* relative permittivity, 3N; +inf marks PEC */
    void* S;          /* Poynting vector, 3N (poynting != OFF) */
    void* oS;         /* 6N partial products (FDTD_POYNTING_TWO_PASS only) */
    void* F;          /* Bloch partner of E (bloch_mask != 0) */
    void* G;          /* Bloch partner of H, read */
    void* G_out;      /* Bloch partner of H, written (same rule as H_out) */
    void* psi_e;      /* CPML psi_E block, fdtd_psi_elems() values, zero-initialised */
    void* psi_h;      /* CPML psi_H block read */
    void* psi_h_out;  /* CPML psi_H block written (== psi_h when H_out == H) */
    void* psi_f;      /* psi blocks of the Bloch partner system */
    void* psi_g;
    void* psi_g_out;
    void* coef;       /* fdtd_coef_elems() values: per face p0e,p1e,p2e,p0h,p1h,p2h (T each) */
    void* inv_d;      /* nonuniform only: 1/dx_e along x,y,z then 1/dx_h along x,y,z,
                         lengths sx,sy,sz,sx,sy,sz (simulation.py:136-142, 179-182) */
    void* E_out;      /* optional second E buffer (3N, != E): written by fdtd_update_he;
                         lets fdtd_step_n fuse each H half-step with the next E half-step.
                         NULL: no fusion */
    void* psi_e_out;  /* optional second psi_E block (!= psi_e): written by fdtd_update_he2;
                         with E_out it lets fdtd_step_n run two steps per pass.  NULL: none */
    void* F_out;      /* optional second F buffer (3N, != F): with Bloch boundaries
                         fdtd_update_he writes F there (with G_out != G, psi_g_out != psi_g).
                         NULL: none */
} fdtd_fields;

int fdtd_abi_version(void);
const char* fdtd_last_error(void);

/* Validates the description and precomputes launch shapes.  Host memory only. */
int fdtd_plan_create(const fdtd_desc* desc, void** plan_out);
int fdtd_plan_destroy(void* plan);

/* Text of the specialisation behind one update (which: 0 update_E, 1
 * update_H, 2 update_S, 3 update_F, 4 update_G, 5 step_n): kernel, template
 * flags, launch shape, CPML faces and Bloch axes.  Replaces the rendered
 * template sources the reference keeps in sim.sources (simulation.py:209-230).
 * Copies at most size-1 chars and a NUL into buf (when buf is not NULL) and
 * returns the full length, or -1 on error. */
int64_t fdtd_plan_describe(const void* plan, int32_t which, char* buf, int64_t size);

/* Sizes (in elements of the field type) of the caller-owned blocks. */
int64_t fdtd_psi_elems(const void* plan);
int64_t fdtd_coef_elems(const void* plan);

/* One E half-step (E, and F when Bloch is on). */
int fdtd_update_e(const void* plan, const fdtd_fields* f, void* stream);
/* One H half-step (H, and G when Bloch is on); writes S when poynting is FUSED,
   oS when it is TWO_PASS. */
int fdtd_update_h(const void* plan, const fdtd_fields* f, void* stream);
/* S from oS (TWO_PASS); a no-op returning 0 for FUSED. */
int fdtd_update_s(const void* plan, const fdtd_fields* f, void* stream);
/* E += c*Jr + s*Ji for xmin<=x<xmax, ymin<=y<ymax, zmin<=z<zmax; box = {xmin,xmax,ymin,ymax,zmin,zmax}. */
int fdtd_update_j(const void* plan, const fdtd_fields* f, const void* Jr, const void* Ji,
                  double c, double s, const int64_t* box, void* stream);
/* field[index] += amp[0]; amp is a device pointer to one value of the field type. */
int fdtd_add_point(const void* plan, void* field, int64_t index, const void* amp, void* stream);

/* n full steps in the order of the reference caller loop (fdtd.py:162-175):
 * E (+F), E[src_index] += src_wave[k] (skipped when src_wave is NULL),
 * H (+G) with S.  With H_out != H the H (and G, psi_h, psi_g) buffers swap
 * roles every step: after an odd n the newest H is in f->H_out. */
int fdtd_step_n(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index,
                const void* src_wave, void* stream);
/* fdtd_step_n (same loop, same results) without bringing E and psi_E back
 * to their own buffers: with f->E_out (and f->psi_e_out) set it runs
 * (n - 1) / 2 two-step passes and at most one one-step pass, where
 * fdtd_step_n keeps the pass counts even.  *moved receives bit 0 when the
 * new E is in f->E_out, bit 1 when psi_E is in f->psi_e_out, bit 2 when the
 * newest H (with psi_H, G, psi_G) is in f->H_out.  The caller swaps its
 * buffer roles accordingly. */
int fdtd_step_n_moved(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index,
                      const void* src_wave, void* stream, int32_t* moved);

/* fdtd_step_n_moved (same loop, same results, same *moved bits) chained
 * over calls.  flags:
 *   FDTD_CHAIN_H_PENDING  the H half-step of the previous call's last step
 *                         has not run: it runs first, in one pass with E_0;
 *   FDTD_CHAIN_DEFER_H    leave the H half-step of the last step (H_{n-1},
 *                         with S) unrun.
 * *moved also receives bit 3 (FDTD_MOVED_H_PENDING) when an H half-step is
 * pending on return (H, psi_H and S then hold the step before); the caller
 * passes FDTD_CHAIN_H_PENDING to the next call or runs fdtd_update_h.
 * Consecutive calls with both flags run every H half-step fused with the
 * next E half-step: n steps in n / 2 two-step passes.  No counterpart in
 * the reference, whose caller loop (fdtd.py:162-175) has no call boundary
 * on the device. */
#define FDTD_CHAIN_H_PENDING 1
#define FDTD_CHAIN_DEFER_H 2
#define FDTD_MOVED_H_PENDING 8
int fdtd_step_n_chain(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index,
                      const void* src_wave, int32_t flags, void* stream, int32_t* moved);

/* What fdtd_step_n_observed records after every step k = 0 .. n-1.  No
 * counterpart in the reference: a caller there reads sim.S (or E, H) on the
 * host between steps; this keeps the stepping on the device for the
 * objective and checkpointed recompute of the adjoint. */
typedef struct fdtd_observe {
    double* flux_acc;        /* NULL: none; else flux_acc[0] += the sum of S component flux_comp over the
                                plane {axis flux_axis coordinate == flux_index} (as fdtd_plane_sum), per step */
    int32_t flux_comp, flux_axis;
    int64_t flux_index;
    double* flux_partials;   /* NULL: each step's sum as fdtd_plane_sum (one workgroup); else
                                fdtd_reduce_partials(plan) doubles of device scratch for the same sum over
                                many workgroups (another fixed order; faster on planes normal to z) */
    void* const* E_keep;     /* NULL, or n device buffers of 3N values: E after step k is copied to E_keep[k] */
    void* const* H_keep;     /* NULL, or the same for H */
} fdtd_observe;

/* fdtd_step_n_moved (same loop, same results, same *moved bits) with the
 * observations of *obs after every step.  It runs one step per pass (the
 * two-step pass does not keep the first step's S); obs may be NULL. */
int fdtd_step_n_observed(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index,
                         const void* src_wave, const fdtd_observe* obs, void* stream, int32_t* moved);

/* The H half-step with S followed by the next E half-step in one pass:
 * H_out = H-step(E, H); S; E_out = E-step(E, H_out); then, when src_amp is
 * not NULL, E_out[src_index] += src_amp[0].  E is left unchanged (E_out holds
 * the new E), psi_e is updated in place, psi_h -> psi_h_out.  Bit-identical to
 * fdtd_update_h, then fdtd_update_e on E_out := E, then fdtd_add_point.
 * Needs fdtd_update_he_supported(plan) (fused Poynting; long rows are cut
 * into z segments) and distinct E/E_out, H/H_out, psi_h/psi_h_out.  With
 * Bloch boundaries (FP32, uniform dx, rows of at most 512 cells, each tile
 * whole rows) the pass advances the pair: F -> F_out and
 * G -> G_out, psi_g -> psi_g_out, psi_f in place, all distinct; it is then
 * bit-identical to the paired fdtd_update_h and fdtd_update_e (the four
 * update_H/G/E/F launches of simulation.py:252-264).
 * With f->E_out set and a supported plan, fdtd_step_n uses it for all but
 * the first E and last H half-steps and arranges the count so that the new E
 * ends in f->E. */
int fdtd_update_he(const void* plan, const fdtd_fields* f, int64_t src_index, const void* src_amp,
                   void* stream);
/* 1 when fdtd_update_he can run with this plan, else 0. */
int fdtd_update_he_supported(const void* plan);

/* Two iterations of the reference caller loop (fdtd.py:162-175) from the
 * middle of a step, in one pass over the grid: H-step (with S) on E, H;
 * E-step; E[src_index] += src_amp[0]; H-step (with S); E-step;
 * E[src_index] += src_amp[1] (src_amp: NULL, or a device pointer to two
 * values).  Reads E, H, psi_e, psi_h and writes E_out, H_out, psi_e_out,
 * psi_h_out, S (S of the second H-step).  Bit-identical to fdtd_update_he
 * twice (or to update_H/update_S (update_h.cl, update_s.cl), update_E
 * (update_e.cl) and the source line of fdtd.py:169, twice).  Needs
 * fdtd_update_he2_supported(plan) and distinct E/E_out, H/H_out,
 * psi_h/psi_h_out, psi_e/psi_e_out.  With f->E_out and f->psi_e_out set,
 * fdtd_step_n runs most of its steps two per pass; psi_e ends in f->psi_e. */
int fdtd_update_he2(const void* plan, const fdtd_fields* f, int64_t src_index, const void* src_amp,
                    void* stream);
/* 1 when fdtd_update_he2 can run with this plan (FP32, fused Poynting, no
   Bloch boundaries, CPML slabs of one axis apart), else 0. */
int fdtd_update_he2_supported(const void* plan);
/* 1 when fdtd_step_n runs two steps per pass with this plan (given E_out and
   psi_e_out): fdtd_update_he2 is supported and its time model (rounds of
   workgroups x planes per chunk) beats the one-step fused pass's, else 0.
   Where the two-step tiles (three halo rows, and three halo columns when
   rows are cut into z segments) own few of their cells, the one-step pass
   is kept. */
int fdtd_update_he2_preferred(const void* plan);

/* ------------------------------------------------------------------------
 * Discrete adjoint of the step and FP64 reductions.  No counterpart in the
 * reference: they serve gradient-based design (d objective / d eps) on top
 * of the same step.  Supported for uniform dx, no Bloch boundaries and
 * non-overlapping CPML slabs.
 *
 * Objective: J = sum over steps of sum over the plane {coordinate of axis
 * obj_axis == obj_index} of S_{obj_axis}.  One adjoint step reverses one
 * forward step (E step, source, H step with S), given the forward fields
 * before and after it.
 * ------------------------------------------------------------------------ */
typedef struct {
    const void* E_prev;   /* E before the step's E update */
    const void* E_next;   /* E after the E update and the source */
    const void* H_prev;   /* H before the step's H update */
    const void* H_next;   /* H after the H update */
    const void* eps;
    void* Eb;             /* adjoint of E: in = after the step, out = before (in place) */
    void* Hb;             /* adjoint of H, same convention */
    const void* psib_e;   /* adjoint of psi_E after the step */
    void* psib_e_out;     /* adjoint of psi_E before the step (distinct buffer) */
    const void* psib_h;
    void* psib_h_out;
    const void* coef;     /* the forward coefficient block */
    double* eps_grad;     /* 3N doubles, accumulated d J / d eps (component-wise), or NULL */
    int64_t src_index;    /* forward point source index, or -1 */
    double src_amp;       /* value added at src_index after the E update */
    int32_t obj_axis;     /* 0..2, or -1 for no objective term this step */
    int32_t obj_index;
    int32_t parts;        /* 0 or 3: the whole step; 1: objective + H step only;
                             2: E step only (lets a caller insert its own
                             adjoint between the two, e.g. the shell net) */
    int32_t reserved;
    int32_t grad_lo[3];   /* eps_grad is accumulated only inside the box grad_lo <= (x, y, z) < grad_hi */
    int32_t grad_hi[3];   /* (a design region); grad_hi all zero: the whole grid */
} fdtd_adjoint_bufs;

/* Elements of one adjoint psi block (same as fdtd_psi_elems). */
int fdtd_adjoint_step(const void* plan, const fdtd_adjoint_bufs* b, void* stream);

/* acc[0] += sum of field component comp over the plane {axis coordinate ==
 * index}, accumulated in FP64 in a fixed order (deterministic). */
int fdtd_plane_sum(const void* plan, const void* field, int32_t comp, int32_t axis, int64_t index,
                   double* acc, void* stream);

/* Electromagnetic energy sum(eps E.E) + sum(H.H_prev) in FP64 (the quantity
 * the leapfrog step conserves without sources and losses); partials must hold
 * fdtd_reduce_partials() doubles; the result is written to out[0]. */
int64_t fdtd_reduce_partials(const void* plan);
int fdtd_energy(const void* plan, const void* E, const void* H, const void* H_prev, const void* eps,
                double* partials, double* out, void* stream);
/* fdtd_energy without the partials argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  It allocates nothing: the partials live in a file-scope
 * device array of the library (1024 doubles, 8 KiB, one copy per GPU, held
 * until the process ends); a plan whose reduction needs more is rejected.
 * Same kernels and summation order as fdtd_energy, so the result is
 * bit-identical.  Calls that overlap on one GPU share that array and race:
 * serialise them (one stream, or events). */
int fdtd_energy_noalloc(const void* plan, const void* E, const void* H, const void* H_prev, const void* eps,
                        double* out, void* stream);

/* ------------------------------------------------------------------------
 * Learned absorbing shell.  No counterpart in the reference: a thin shell of
 * T cells on every face runs the CPML-T update, and a linear 2-layer 3x3x3
 * convolution (6 -> 16 -> 3 channels: E, H in, a correction to E out) of
 * the fields after the E update (and source) is added to E on the shell
 * cells every step, every second or every fourth step (fdtd_shell_step_n's
 * period).  Six
 * weight sets, one per face box, each with one layer-2 head per depth
 * (index along the box's thin axis); BF16 weights.  Inference evaluates the
 * two layers as their composed 5x5x5 convolution (composed in FP32, rounded
 * once to BF16) on BF16 inputs with FP32 accumulation.
 * Weights: fdtd_shell_weight_elems() BF16 values, W1[box][16][6][27] for the
 * six boxes, then W2[box][depth 0..3][3][16][27]; taps
 * t = 9(dx+1) + 3(dy+1) + (dz+1).
 * ------------------------------------------------------------------------ */
int64_t fdtd_shell_weight_elems(void);
/* FP32 scratch of the correction (the hidden layer or the correction
 * itself, then the weights repacked in matrix-core fragment order); the
 * backward also takes a second buffer of this size for the hidden adjoint. */
int64_t fdtd_shell_hidden_elems(const void* plan, int32_t thickness);
/* FP32 scratch for the weight-gradient partials. */
int64_t fdtd_shell_partials_elems(void);
/* E += correction(E, H) on the shell cells (E read before it is corrected). */
int fdtd_shell_apply(const void* plan, int32_t thickness, const uint16_t* w, void* E, const void* H,
                     float* hidden, void* stream);
/* n steps of the shell solver.  Step k: E update (with the pending
 * correction added to E first) and source, as fdtd_step_n; H update from
 * that E; then, when the step is corrected, E += correction(E, H) with E
 * and H as the H update read them.  period 1: every step is corrected (the
 * addition is made by the next step's E update, and by a final pass after
 * step n); period 2 or 4: steps 0, period, 2 period, ... of the call (n a
 * multiple of the period; each correction is added by the next step's E
 * update).  With period 2 or 4, FP32, a uniform grid, the fused passes'
 * buffers and moved non-NULL, each pair of steps runs as one two-step pass
 * with the corrections between passes, and E, psi_E and H may end in the
 * other buffers: *moved gets the bits of fdtd_step_n_moved.  Otherwise E and psi_E end in f->E and f->psi_e (H as
 * fdtd_step_n; *moved reports it when moved is not NULL).  Real fields only
 * (no Bloch boundaries), two-pass Poynting excluded. */
int fdtd_shell_step_n(const void* plan, const fdtd_fields* f, int32_t thickness, const uint16_t* w,
                      float* hidden, int32_t n, int32_t period, int64_t src_index, const void* src_wave,
                      void* stream, int32_t* moved);
/* Adjoint of E_out = E + correction(E, H): given Ebar (adjoint of E_out),
 * Ebar_out += Ebar-contribution through the correction's E input (callers
 * initialise Ebar_out with Ebar for the identity part), Hbar_out += the H
 * part, and wgrad (FP32, weight layout) += the weight gradient.  E and H are
 * the correction's inputs. */
int fdtd_shell_backward(const void* plan, int32_t thickness, const uint16_t* w, const void* E, const void* H,
                        float* hidden, float* hbar, const void* Ebar, void* Ebar_out, void* Hbar_out,
                        float* partials, float* wgrad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_MI355X_H */
