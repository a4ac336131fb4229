/*
 * frfs_fq.h -- C ABI of the MI355X fast spectral collision operator Q(f,f).
 *
 * Drop-in boundary for the reference's scattering-model plugin call.  The
 * reference evaluates Q one spatial DOF at a time from Python:
 *
 *   DGFSVHSGLLScatteringModel.fs(d_arr_in, d_arr_out, elem, upt)
 *       frfs/solvers/dgfs/scattering.py:227-286      (vhs-gll)
 *   DGFSVHSGLLNodalScatteringModel.fs(f, g, Q, elem, upt)
 *       frfs/solvers/dgfs/scattering.py:327-414      (vhs-gll-nodal)
 *   DGFSBiVSSGLLScatteringModel.fs(elem, upt, *[f_0..f_S-1, Q_0..Q_S-1])
 *       frfs/solvers/dgfsbi/scattering.py:323-330    (vss-gll)
 *
 * each issuing ~8 PyCUDA kernel launches and 6-7 cuFFT executions
 * (frfs/solvers/dgfs/cufft.py:231-281) on the legacy default stream.  This
 * library replaces the Mako/PyCUDA kernels, cuFFT and the per-DOF loop of
 * DGFSSystem.collide (frfs/solvers/dgfs/system.py:199-218) with batched HIP
 * kernels: one fq_eval call evaluates Q for a contiguous range of DOFs.
 *
 * Data boundary: f and Q are the solver's AoSoA matrices of ioshape
 * (nupts, N^3, neles) (frfs/backends/base/types.py:26-47), addressed as
 *   upt*ldim + ((elem/soasz)*N^3 + v)*soasz + elem%soasz
 * (SOA_IX, frfs/backends/cuda/kernels/base.mako:5-6), with DOF index
 * dof = elem*nupts + upt (the reference's ndrange(neles, nupts) order).
 *
 * Errors: functions return FQ_SUCCESS (0) or a negative FQ_ERR_* code and
 * set a thread-local message readable with fq_last_error(); the Python shim
 * raises from it, as frfs/solvers/dgfs/cufft.py:48-117 does for cuFFT.
 *
 * Ownership: the caller owns f and Q device memory; the handle owns the
 * precomputed tables and the scratch.  One handle per stream: calls on one
 * handle must not overlap (they share its scratch).
 */
#ifndef FRFS_FQ_H
#define FRFS_FQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FQ_ABI_VERSION 1

enum fq_status {
    FQ_SUCCESS = 0,
    FQ_ERR_INVALID = -1,      /* bad argument or parameter */
    FQ_ERR_UNSUPPORTED = -2,  /* valid but not built (e.g. grid size) */
    FQ_ERR_HIP = -3,          /* HIP runtime error */
    FQ_ERR_ALLOC = -4         /* device allocation failed */
};

enum fq_model {
    FQ_MODEL_VHS = 0,         /* vhs-gll: half-sphere, alpha = 1 */
    FQ_MODEL_VHS_NODAL = 1,   /* vhs-gll-nodal: bilinear Q(f,g) */
    FQ_MODEL_VSS = 2          /* vss-gll: full sphere, nspcs species */
};

enum fq_precision {
    FQ_PREC_FP64 = 0,         /* every stage in FP64 (the reference's arithmetic) */
    FQ_PREC_MIXED = 1         /* FP32 transforms/products, FP64 accumulation */
};

enum fq_flags {
    FQ_FLAG_PROJECT = 1,      /* least-squares FP64 correction of Q that zeroes the
                                 collision invariants: the mass of every species
                                 and the total (mass-weighted) momentum and energy */
    FQ_FLAG_UNFUSED = 2,      /* stage every inverse transform through HBM
                                 (three-kernel path; for comparison) */
    FQ_FLAG_TWO_LANE = 4,     /* FP32 on 24^3/32^3 grids: product kernel with two
                                 lanes per pencil instead of one (a schedule
                                 candidate of the autotuner) */
    FQ_FLAG_NO_PACK = 8,      /* FP32 one-lane product kernel: transform the two
                                 items of a product separately instead of as one
                                 packed complex volume with Nyquist-plane
                                 corrections (for comparison) */
    FQ_FLAG_NO_SCRATCH = 16,  /* fq_create allocates the tables only, no scratch:
                                 the handle evaluates through fq_eval_noalloc */
    FQ_FLAG_ACCUM_QUAD = 32,  /* 32^3 grid, FP32 transforms: the accumulation
                                 kernel shares each weight row over four DOFs per
                                 workgroup (two workgroups per CU) instead of
                                 eight (one per CU); a schedule candidate of the
                                 autotuner */
    FQ_FLAG_NO_HPAIR = 64,    /* 32^3 grid, FP32 transforms: write one product
                                 volume per group instead of one per group pair
                                 (the two groups' real parts as one complex
                                 volume, the Hermitian part of the gain spectrum
                                 accumulated on half the rows); for comparison */
    FQ_FLAG_OVERLAP = 128     /* two scratch sets: consecutive internal batches
                                 alternate between them, the odd ones on a second
                                 stream of the handle (forked from and joined
                                 back to the caller's stream with events), so
                                 one batch's memory-bound stages can run beside
                                 the next batch's product kernel */,
    FQ_FLAG_LIBRARY_FFT = 256 /* every 3D transform in the ROCm FFT library
                                 (hipFFT) with pointwise kernels between them,
                                 as on grids without native kernels (any even
                                 Nv from 4 to 128 other than 8, 16, 24, 32);
                                 for comparison on the native grids */
};

/* Size of the static device scratch of fq_eval_noalloc (one copy per GPU the
 * library runs on, held until the process ends). */
#define FQ_NOALLOC_SCRATCH_BYTES (256ll << 20)

/*
 * Model parameters.  The host builds the velocity mesh exactly as the
 * reference does (frfs/solvers/dgfs/velocitymesh.py:25-135,
 * frfs/solvers/dgfsbi/velocitymesh.py:27-151) and passes the results.
 */
typedef struct fq_params {
    int32_t abi_version;      /* FQ_ABI_VERSION */
    int32_t model;            /* enum fq_model */
    int32_t precision;        /* enum fq_precision */
    int32_t flags;            /* enum fq_flags */
    int32_t Nv;               /* velocity points per axis (even) */
    int32_t Nrho;             /* radial Gauss points */
    int32_t M;                /* spherical directions used (half sphere for VHS) */
    int32_t nspcs;            /* species (1 for VHS) */
    int32_t Mpre;             /* VSS pre-integration design points */
    int32_t soasz;            /* AoSoA block size (32 in the reference) */
    int64_t max_chunk;        /* DOFs per internal batch; 0 = library default */
    double L;                 /* velocity box half-width */
    double sw;                /* spherical weight: 2 pi/M (VHS) or 4 pi/M (VSS) */
    double cw;                /* velocity cell volume (2L/N)^3, for the projection */
    const double *qz;         /* [Nrho] radial nodes on [0, R] */
    const double *qw;         /* [Nrho] radial weights */
    const double *sz;         /* [M*3] unit directions, row major */
    /* VHS / VHS_NODAL */
    double gamma;             /* 2(1 - omega) */
    double prefac;            /* collision prefactor */
    /* VSS: pair index pq = p*nspcs + q */
    const double *masses;     /* [nspcs] */
    const double *eta;        /* [nspcs^2] alpha_pq - 1 */
    const double *gammapq;    /* [nspcs^2] */
    const double *prefacpq;   /* [nspcs^2] */
    const double *szpre;      /* [Mpre*3] */
    double swpre;             /* 4 pi/Mpre */
} fq_params;

typedef struct fq_handle fq_handle;
typedef struct fq_plan fq_plan;

/* Plan summary: work counts of one DOF-evaluation. */
typedef struct fq_plan_info {
    int32_t n_items;          /* distinct phase-shifted inverse 3D transforms */
    int32_t n_products;       /* pointwise products E1*E2 */
    int32_t n_groups;         /* forward 3D transforms (product sums) */
    int32_t n_contribs;       /* weighted accumulations into Q-hat */
    int32_t n_loss;           /* loss-term tables */
    int32_t n_sources;        /* input distributions */
    int32_t n_targets;        /* output distributions */
    int32_t n_fft_ref;        /* 3D transforms per DOF in the reference */
    int32_t n_fft;            /* 3D transforms per DOF here */
} fq_plan_info;

/* ---- lifecycle -------------------------------------------------------- */

/* Replaces DGFSScatteringModel.__init__ -> load_parameters() ->
 * perform_precomputation() (frfs/solvers/dgfs/scattering.py:17-28,67-224;
 * frfs/solvers/dgfsbi/scattering.py:93-256): builds the transform plan,
 * uploads the phase tables and computes the kernel weight tables on the
 * current HIP device (precompute_aa/_bb, vhs-gll.mako:45-101;
 * precompute_a/_bc, vss-gll.mako:31-95). */
int fq_create(const fq_params *prm, fq_handle **out);
int fq_destroy(fq_handle *h);
const char *fq_last_error(void);
int fq_get_plan_info(const fq_handle *h, fq_plan_info *info);
/* Bytes of device scratch held by the handle. */
int fq_scratch_bytes(const fq_handle *h, int64_t *bytes);

/* ---- evaluation (device pointers; stream is a hipStream_t, NULL = default) */

/* Q for DOFs [first_dof, first_dof+ndof): replaces vhs-gll fs
 * (frfs/solvers/dgfs/scattering.py:227-286) and single-species vss-gll fs
 * (frfs/solvers/dgfsbi/scattering.py:323-330), and the collide loop around
 * them (frfs/solvers/dgfs/system.py:217-218). */
int fq_eval(fq_handle *h, const double *f, double *q,
            int32_t nupts, int64_t ldim, int32_t neles,
            int64_t first_dof, int64_t ndof, void *stream);

/* fq_eval without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and results
 * as fq_eval (the same kernels, bit for bit), but the batch scratch lives in
 * a static device array of FQ_NOALLOC_SCRATCH_BYTES per GPU instead of the
 * handle's scratch: the DOF range is evaluated in batches of as many DOFs as
 * fit there.  Returns FQ_ERR_UNSUPPORTED when one DOF of the plan needs
 * more than FQ_NOALLOC_SCRATCH_BYTES (e.g. a 32^3 two-species FP64 plan is
 * not served; 32^3 single-species mixed takes 13 DOFs per batch).  Works on
 * any single-species handle, including one created with
 * FQ_FLAG_NO_SCRATCH.  Every call on a GPU shares that GPU's array, which
 * keeps the last call's contents (every kernel writes what it later reads,
 * so nothing needs clearing): calls that overlap on one GPU, on any handle
 * or stream, race. */
int fq_eval_noalloc(fq_handle *h, const double *f, double *q,
                    int32_t nupts, int64_t ldim, int32_t neles,
                    int64_t first_dof, int64_t ndof, void *stream);

/* Multi-species vss-gll: f[s], q[s] for s < nspcs, all with the same AoSoA
 * geometry (frfs/solvers/dgfsbi/scattering.py:259-330;
 * frfs/solvers/dgfsbi/system.py:239-251). */
int fq_eval_multi(fq_handle *h, const double *const *f, double *const *q,
                  int32_t nupts, int64_t ldim, int32_t neles,
                  int64_t first_dof, int64_t ndof, void *stream);

/* Bilinear vhs-gll-nodal Q(f, g) (frfs/solvers/dgfs/scattering.py:327-414). */
int fq_eval_bilinear(fq_handle *h, const double *f, const double *g, double *q,
                     int32_t nupts, int64_t ldim, int32_t neles,
                     int64_t first_dof, int64_t ndof, void *stream);

/* ---- instrumentation --------------------------------------------------- */

#define FQ_NKERNELS 10
/* Bracket every kernel launch with HIP events on the launch stream (on = 1)
 * or stop (on = 0); either call discards earlier records.  No reference
 * counterpart: the reference times only its precompute with cuda.Event
 * (frfs/solvers/dgfsbi/scattering.py:166-205). */
int fq_profile_enable(fq_handle *h, int on);
/* Wait for the recorded events; ms[k] / launches[k] = summed duration and
 * launch count of kernel k (k < FQ_NKERNELS) since the last read. */
int fq_profile_read(fq_handle *h, double *ms, int64_t *launches);
const char *fq_kernel_name(int k);
/* DOFs per internal batch (kernel launch sequence). */
int fq_chunk_dofs(const fq_handle *h, int64_t *chunk);
/* The kernel paths this handle runs (a bitmask of FQ_PATH_*), chosen at
 * fq_create from the grid, precision and flags.  No reference counterpart. */
enum fq_path {
    FQ_PATH_FUSED = 1,        /* product kernel with the inverse z stage inside */
    FQ_PATH_ONE_LANE = 2,     /* FP32 product kernel with one lane per pencil */
    FQ_PATH_PAIR_PACK = 4,    /* the two items of a product as one complex volume */
    FQ_PATH_GROUP_PAIRS = 8,  /* two groups' real parts as one product volume */
    FQ_PATH_OVERLAP = 16,     /* two scratch sets and two streams */
    FQ_PATH_LIBRARY_FFT = 32  /* grid without native kernels: hipFFT transforms */
};
int fq_get_paths(const fq_handle *h, int32_t *paths);

/* ---- kernels on either side of Q (stateless, AoSoA banks) -------------- */

#define FQ_AXNPBY_MAX 4
/* x0[i] = a[0]*x0[i] + sum_{k<nv} a[k+1]*xs[k][i] for i < n over whole banks
 * (a[0] == 0 never reads x0; a[0] == 1 accumulates): replaces the
 * axnpby_dgfs_full kernel (frfs/backends/cuda/kernels/axnpby_dgfs_full.mako,
 * frfs/backends/cuda/blasext.py:46-75).  nv <= FQ_AXNPBY_MAX. */
int fq_axnpby(int64_t n, double *x0, int32_t nv, const double *const *xs, const double *a, void *stream);

/* dst[u, sidx_dst + v, e] = src[u, sidx_src + v, e] for v < count, every
 * point u < nrow and element e < neles, between AoSoA matrices of nvar_src /
 * nvar_dst variables: replaces copy_dgfs (frfs/backends/cuda/kernels/copy_dgfs.mako,
 * frfs/backends/cuda/blasext.py:77-138, copy_to_reg / copy_from_reg). */
int fq_copy_dgfs(int32_t nrow, int32_t neles, int32_t count, const double *src, int64_t ldim_src,
                 int32_t nvar_src, int32_t sidx_src, double *dst, int64_t ldim_dst, int32_t nvar_dst,
                 int32_t sidx_dst, void *stream);

#define FQ_NMOMENTS 14
/* Non-dimensional moments per DOF, out[d*FQ_NMOMENTS + k] in the order of
 * _compute_moments_3D (frfs/plugins/dgfsmomwriterstd.py:142-214):
 * rho, ux, uy, uz, T, qx, qy, qz, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz, for the
 * velocity grid of Nv points per axis on [-L, L]^3 with cell volume cw and
 * species mass ratio mass_ratio. */
int fq_moments(int32_t Nv, double L, double cw, double mass_ratio, const double *f, int32_t nupts,
               int64_t ldim, int32_t neles, int64_t first_dof, int64_t ndof, double *out, void *stream);

/* Flux-reconstruction advection right-hand side out = -div(c f) for every
 * velocity v < nvar of AoSoA banks of ioshape ((order+1)^2, nvar, nelx*nely)
 * on a periodic structured mesh of affine hx x hy quads (element
 * e = ex + nelx*ey), with upwind common fluxes: replaces the reference's
 * tflux / intcflux / rusanov / negdivconf kernels and the M0-M3 operator
 * products (frfs/solvers/dgfs/kernels/{tflux,intcflux,negdivconf}.mako,
 * rsolvers/rusanov.mako, frfs/solvers/dgfs/elements.py:26-164) for this mesh
 * class.  ops = {Dx, Dy, M0, M3} from frfs_amd/fr.py (device, row major);
 * cvx/cvy = velocity components (device, nvar).  order in {1, 2, 3}. */
int fq_fr_rhs_quad(int32_t order, int32_t nelx, int32_t nely, double hx, double hy, const double *ops,
                   const double *cvx, const double *cvy, const double *u, double *out, int64_t ldim,
                   int32_t nvar, void *stream);

/* ---- host-only plan inspection (no GPU needed) ------------------------- */

int fq_plan_create(const fq_params *prm, fq_plan **out);
int fq_plan_destroy(fq_plan *p);
int fq_plan_get_info(const fq_plan *p, fq_plan_info *info);
/* Export the plan arrays (any pointer may be NULL):
 *   item_src[n_items], item_theta[n_items*3],
 *   prod_i1/prod_i2[n_products], prod_coef[n_products], prod_group[n_products],
 *   contrib_group/contrib_target[n_contribs],
 *   loss_src/loss_target[n_loss]. */
int fq_plan_export(const fq_plan *p, int32_t *item_src, double *item_theta,
                   int32_t *prod_i1, int32_t *prod_i2, double *prod_coef,
                   int32_t *prod_group, int32_t *contrib_group,
                   int32_t *contrib_target, int32_t *loss_src,
                   int32_t *loss_target);

#ifdef __cplusplus
}
#endif

#endif /* FRFS_FQ_H */
