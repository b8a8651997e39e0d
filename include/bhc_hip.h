/*
 * bhc_hip.h -- C ABI of the MI355X-native BELLHOP ray/beam tracing hot path.
 *
 * Two layers are exported:
 *
 *  1. Preprocessing (host C++): a raw 2D environment description (the values an
 *     .env/.bty/.ati/.brc file holds, in user units) is turned into the
 *     preprocessed problem that the reference hands to its launch layer.
 *     Replaces the Preprocess() step of the reference's params modules,
 *     src/api.cpp:391-393 -> src/module/{ssp.hpp:453-545, boundary.hpp:499-544,
 *     boundarycond.hpp:166-224, atten.cpp:99-194, beaminfo.hpp:221-238,
 *     rayangles.hpp:157-190, rcvrranges.hpp:74-85, szrz.hpp:138-142, sbp.hpp:143-153}.
 *
 *  2. The device boundary (HIP): a preprocessed problem plus caller-owned output
 *     buffers in, accumulated field / arrivals / eigen hits / ray paths out.
 *     Replaces bhc::mode::RunFieldModesImpl<CFG,O3D,R3D>(bhcParams&, bhcOutputs&)
 *     (src/mode/fieldimpl.hpp:27-28, CPU src/mode/fieldimpl.cpp.in:55-77,
 *     CUDA src/mode/fieldimpl.cu.in:51-64) and the ray-mode worker pool
 *     RunRayMode (src/mode/ray.cpp:167-188).
 *
 * All real-valued scalars and arrays cross the ABI as double. In 32-bit mode
 * (precision == 32) every such value holds a float exactly (preprocessing is
 * done in float, as the reference's BHC_USE_FLOATS build does), so the
 * conversion is lossless both ways. Receiver and source depths are float in
 * both modes, as in the reference (Position::Rz, Position::Sz).
 *
 * Every entry point returns an int status: 0 = success, < 0 = failure (message
 * via bhc_last_error()). Nothing throws across the ABI and nothing resets the
 * device. Physics errors and warnings are reported through err/warn bitmasks
 * that use the reference numbering (src/util/errors.hpp:61-93).
 */
#ifndef BHC_HIP_H
#define BHC_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHC_ABI_VERSION 1

/* ---- error / warning bit numbers (src/util/errors.hpp:61-93) ---------------- */
#define BHC_ERR_TEMPLATE 0
#define BHC_ERR_JOBNUM 1
#define BHC_ERR_RAYINIT 2
#define BHC_ERR_BOUNDARY_SEG_CONTAINS_NAN 3
#define BHC_ERR_BOUNDARY_CONDITION_TYPE 4
#define BHC_ERR_INVALID_SSP_TYPE 5
#define BHC_ERR_OUTSIDE_ALTIMETRY 6
#define BHC_ERR_OUTSIDE_BATHYMETRY 7
#define BHC_ERR_OUTSIDE_SSP 8
#define BHC_ERR_QUAD_ISEG 9
#define BHC_ERR_INVALID_IMAGE_INDEX 10

#define BHC_WARN_RAYS_OUTOFMEMORY 0
#define BHC_WARN_ONERAY_OUTOFMEMORY 1
#define BHC_WARN_UNBOUNDED_BEAM 2
#define BHC_WARN_TOO_FEW_BEAMS 3
#define BHC_WARN_SOURCE_OUTSIDE_BOUNDARIES 4
#define BHC_WARN_OUTSIDE_REFLCOEF 5
#define BHC_WARN_OUTSIDE_ALTIMETRY 6
#define BHC_WARN_OUTSIDE_BATHYMETRY 7
#define BHC_WARN_STEP_NEGATIVE_H 8
#define BHC_WARN_TRIDIAG_H_NEGATIVE 9
#define BHC_WARN_TRIDIAG_H_GROWING 10
#define BHC_WARN_CPCHIP_INVALIDXT 16
#define BHC_WARN_CPCHIP_INVALIDCCOEF 17
#define BHC_WARN_OCEANTORAYX_GAVEUP 18

/* ---- output record layouts (byte-identical to the reference) ----------------- */

/* Arrival, include/bhc/structs.hpp:376-380 (40 bytes). */
typedef struct bhc_arrival {
    int32_t NTopBnc, NBotBnc;
    float SrcDeclAngle, SrcAzimAngle, RcvrDeclAngle, RcvrAzimAngle, a, Phase;
    float delay[2]; /* complex delay (re, im) */
} bhc_arrival;

/* EigenHit, include/bhc/structs.hpp:357-365 (36 bytes). */
typedef struct bhc_eigen_hit {
    int32_t is, itheta, ir, iz, isx, isy, isz, ialpha, ibeta;
} bhc_eigen_hit;

/* rayPt<false> in the 64-bit layout, include/bhc/structs.hpp:397-424 (112 bytes). */
typedef struct bhc_ray_pt2d {
    double p[2], q[2];
    int32_t NumTopBnc, NumBotBnc;
    double x[2], t[2];
    double c, Amp, Phase;
    double tau[2];
} bhc_ray_pt2d;

/* rayPt<true> in the 64-bit layout, include/bhc/structs.hpp:397-424 (176 bytes):
 * p, q are 2x2 matrices stored column-major (p[0] = p[0][0], p[1] = p[0][1], ...). */
typedef struct bhc_ray_pt3d {
    double p[4], q[4];
    double phi, dummy;
    int32_t NumTopBnc, NumBotBnc;
    double x[3], t[3];
    double c, Amp, Phase;
    double tau[2];
} bhc_ray_pt3d;

/* ---- raw (user-unit) 2D environment: input to preprocessing ------------------ */

/* Halfspace / boundary condition, HSInfo + HSExtra (include/bhc/structs.hpp:112-125).
 * bc: 'V' vacuum, 'R' rigid, 'A' acousto-elastic halfspace, 'F' reflection table,
 * 'G' grain size. Speeds in m/s, attenuations in the SSP's AttenUnit (halfspace)
 * or dB/wavelength ('L' provinces). */
typedef struct bhc_halfspace_in {
    char bc;
    char pad[7];
    double depth;           /* boundary depth (m): Top = first SSP depth, Bot = last */
    double alphaR, betaR, rho, alphaI, betaI;
    double zTemp;           /* depth used for the crci call ('A') */
    double Mz;              /* grain size ('G') */
} bhc_halfspace_in;

/* Altimetry / bathymetry (BdryInfoTopBot<false>), user units: range in km, depth in m.
 * npts == 0 selects the flat default boundary at halfspace depth. The reference
 * extends a file boundary by one point at each end (src/module/boundary.hpp:203-226);
 * the caller passes only the file's points. */
typedef struct bhc_bdry_in {
    char type[2];           /* [0]: 'L' linear / 'C' curvilinear; [1]: 'S' short / 'L' long */
    char pad[6];
    int32_t npts;
    int32_t pad2;
    const double *r_km;     /* [npts] */
    const double *z;        /* [npts] */
    /* per-point geoacoustics when type[1] == 'L' (each [npts], may be NULL otherwise) */
    const double *alphaR, *betaR, *rho, *alphaI, *betaI;
} bhc_bdry_in;

/* Reflection coefficient table (ReflectionInfoTopBot): theta (deg), |R|, phase (deg). */
typedef struct bhc_refl_in {
    int32_t npts;
    int32_t pad;
    const double *theta, *r, *phi;
} bhc_refl_in;

typedef struct bhc_env2d {
    int32_t precision;      /* 64 or 32 */
    int32_t pad0;
    char title[80];
    double freq0;           /* Hz */
    /* sound speed profile (SSPStructure) */
    char ssp_type;          /* 'N','C','S','P','Q','A' */
    char atten_unit[2];     /* e.g. "W ", "WT", "F " */
    char pad1[5];
    int32_t ssp_npts;
    int32_t ssp_nr;         /* 'Q' only: number of profile ranges */
    const double *ssp_z, *ssp_alphaR, *ssp_alphaI, *ssp_rho, *ssp_betaR, *ssp_betaI;
    const double *ssp_r_km;  /* 'Q' only: [ssp_nr] profile ranges */
    const double *ssp_cmat;  /* 'Q' only: [ssp_npts][ssp_nr], row = depth */
    /* Thorp / Francois-Garrison volume attenuation parameters (AttenInfo) */
    double fg_T, fg_S, fg_pH, fg_z_bar;
    /* boundaries */
    bhc_halfspace_in top, bot;
    bhc_bdry_in ati, bty;
    bhc_refl_in trc, brc;   /* top / bottom reflection tables ('F') */
    /* source / receivers (Position) */
    int32_t NSz, NRz, NRr, pad2;
    const float *Sz;        /* [NSz] m */
    const float *Rz;        /* [NRz] m */
    const double *Rr_km;    /* [NRr] km */
    /* launch angles (AngleInfo alpha), degrees, sorted increasing */
    int32_t Nalpha, iSingle_alpha;
    const double *alpha_deg;
    /* beam (BeamStructure<false>) */
    char RunType[7];        /* [0] C/S/I/A/a/E/R/V/v, [1] beam type G/g/B/b/^, [3] R/X, [4] R/I */
    char BeamType[3];       /* Beam->Type[1..3] (width, curvature, shift); "   " if absent */
    char pad3[6];
    double deltas;          /* m, 0 = automatic */
    double box_r_km, box_z; /* beam box */
    double epsMultiplier, rLoop;
    int32_t iBeamWindow, Nimage;
    /* source beam pattern (SBPInfo), angle (deg) / level (dB) pairs; 0 = omni */
    int32_t NSBPPts, pad4;
    const double *sbp;      /* [2*NSBPPts] */
    /* biological volume attenuation (atten_unit[1] == 'B', AttenInfo::bio):
     * NBioLayers rows of z1, z2 (m), f0 (Hz), Q, a0 (dB/km) */
    int32_t NBioLayers, pad5;
    const double *bio;      /* [5*NBioLayers] */
} bhc_env2d;

/* ---- preprocessed 2D problem: what the launch layer consumes ----------------- */

/* One boundary node (BdryPtFull<false> minus the parse-only fields). */
typedef struct bhc_bdry_node2d {
    double x[2], t[2], n[2], Noden[2], Nodet[2];
    double Len, kappa;
    double cP[2], cS[2], rho;   /* node geoacoustics ('L' boundaries) */
} bhc_bdry_node2d;

typedef struct bhc_hs2d {
    char bc;
    char pad[7];
    double cP[2], cS[2];
    double rho, depth;
} bhc_hs2d;

typedef struct bhc_problem2d {
    int32_t precision;      /* 64 or 32 */
    char RunType[7];
    char BeamType[4];       /* Beam->Type[0..3] */
    char ssp_type;
    int32_t pad0;
    double freq0;
    /* SSP */
    int32_t ssp_npts, ssp_nr;
    const double *ssp_z;     /* [npts] */
    const double *ssp_c;     /* [npts][2] complex */
    const double *ssp_cz;    /* [npts][2] */
    const double *ssp_n2;    /* [npts][2]  ('N') */
    const double *ssp_n2z;   /* [npts][2]  ('N') */
    const double *ssp_coef;  /* [4][npts][2]: cSpline ('S') or cCoef ('P') */
    const double *ssp_rho;   /* [npts] */
    const double *ssp_seg_r; /* [nr] m ('Q') */
    const double *ssp_cmat;  /* [npts][nr] ('Q') */
    const double *ssp_czmat; /* [npts-1][nr] ('Q') */
    /* boundaries */
    bhc_hs2d top_hs, bot_hs;
    char ati_type[2], bty_type[2];
    int32_t ati_npts, bty_npts;
    const bhc_bdry_node2d *ati, *bty;
    int32_t trc_npts, brc_npts;
    const double *trc, *brc;    /* [npts][3] theta (rad->deg as reference: degrees), r, phi (rad) */
    /* positions */
    int32_t NSz, NRz, NRr, NRz_per_range;
    const float *Sz, *Rz;
    const double *Rr;           /* m */
    double Delta_r;
    /* angles */
    int32_t Nalpha, iSingle_alpha;
    const double *alpha;        /* rad */
    const double *alpha_cos;    /* cos(alpha) from host libm */
    const double *alpha_sin;    /* sin(alpha) from host libm */
    double Dalpha;
    /* beam */
    double deltas, box_r, box_z, epsMultiplier, rLoop;
    int32_t iBeamWindow, Nimage;
    int32_t NSBPPts, pad1;
    const double *sbp;          /* [NSBPPts][2] degrees, linear level */
} bhc_problem2d;

/* ---- outputs ----------------------------------------------------------------- */

typedef struct bhc_field_out {
    /* TL (RunType C/S/I): complex field, caller-zeroed, layout
     * [NSz][Ntheta=1][NRz_per_range][NRr] of float pairs (uAllSources, pre-ScalePressure). */
    float *field;
    /* Arrivals (A/a/V/v): [NSz*NRz_per_range*NRr][MaxNArr] and counts. */
    bhc_arrival *arr;
    int32_t *narr;
    int32_t max_narr;
    int32_t allow_merging;  /* 1 reproduces the single-threaded merge order deterministically */
    /* Eigen hits (E/V/v) */
    bhc_eigen_hit *hits;
    int32_t *neigen;
    int32_t eigen_memsize;
    int32_t pad;
    /* Optional per-ray statistics, [number of rays] each, may be NULL:
     * steps taken, top/bottom bounces, final (r, z). */
    int32_t *ray_nsteps, *ray_ntop, *ray_nbot;
    double *ray_xend;       /* [rays][2] */
    /* Error reporting (ErrState) */
    uint32_t err, warn, err_count, warn_count;
    /* Work counters (for the roofline): total steps, contributions */
    int64_t total_steps, total_contrib;
} bhc_field_out;

typedef struct bhc_ray_out {
    bhc_ray_pt2d *points;   /* [NRays][MaxPointsPerRay] */
    int32_t *nsteps;        /* [NRays] */
    int32_t max_points_per_ray;
    int32_t pad;
    uint32_t err, warn, err_count, warn_count;
} bhc_ray_out;

/* ---- preprocessing (host C++) ----------------------------------------------- */

typedef struct bhc_problem_owner bhc_problem_owner;

/* Preprocess a raw environment. On success *out owns all arrays; the view is
 * valid until bhc_problem_free(). */
int bhc_preprocess2d(const bhc_env2d *env, bhc_problem_owner **out);
/* SubTab (src/common_setup.hpp:282-316): fills out[0..n) with
 * fma(i, (x1 - x0) / (n - 1), x0) in double (precision 64) or float (32). */
int bhc_subtab(double x0, double x1, int32_t n, int32_t precision, void *out);
const bhc_problem2d *bhc_problem_view(const bhc_problem_owner *p);
void bhc_problem_free(bhc_problem_owner *p);

/* ---- device context ---------------------------------------------------------- */

typedef struct bhc_ctx bhc_ctx;

int bhc_ctx_create(int device, bhc_ctx **out);
void bhc_ctx_destroy(bhc_ctx *ctx);
const char *bhc_last_error(void);
int bhc_device_info(int device, char *name, int name_len, int *num_cus, int *lds_bytes);

/* One-shot field run (TL / arrivals / eigen): upload, launch, wait, copy back.
 * The drop-in for RunFieldModesImpl. */
int bhc_field_run2d(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_field_out *out);

/* Ray run type: per-ray trajectories, replaces RunRayMode. */
int bhc_ray_run2d(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_ray_out *out);
/* Ray mode over an explicit ray list, for the eigenray re-trace
 * (EigenModePostWorker, src/mode/eigen.cpp:24-48): launch angle and source
 * index per ray and an optional step limit per ray (EigenHit::is). */
int bhc_ray_run2d_list(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_ray_out *out, const int32_t *ialpha,
                       const int32_t *isz, const int32_t *nsteps_in, int32_t nrays);

/* Resident-problem interface (inputs stay in HBM across launches). `spec`
 * gives the output shapes: max_narr / allow_merging (arrivals), eigen_memsize
 * (eigen hits), and whether per-ray statistics are kept (ray_nsteps != NULL);
 * its pointers are not retained. */
typedef struct bhc_dev_problem bhc_dev_problem;
int bhc_problem_upload2d(bhc_ctx *ctx, const bhc_problem2d *prob, const bhc_field_out *spec,
                         bhc_dev_problem **out);
int bhc_field_launch(bhc_ctx *ctx, bhc_dev_problem *dp);   /* zero field + trace, async */
/* bhc_field_launch without any allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory during a run (serving
 * loops, code under HIP graph capture, harnesses that forbid hipMalloc). The
 * trace -> scatter segment buffer is a static 64 MiB device array instead of
 * the context's (allocated at its first use); rays are split over as many
 * passes as that takes, and the scatter uses the wavefront / lane mappings
 * (no range bins). TL run types (C, S, I) only; the problem must have been
 * uploaded (bhc_problem_upload2d). Rejected with an error if one pass could
 * not hold a ray's segments. Calls on one GPU share its array: calls that
 * overlap on one GPU (from several contexts or threads) race. */
int bhc_field_launch_noalloc(bhc_ctx *ctx, bhc_dev_problem *dp);
int bhc_field_download(bhc_ctx *ctx, bhc_dev_problem *dp, bhc_field_out *out);
void bhc_problem_release(bhc_dev_problem *dp);
/* Scatter-mapping search for a resident problem (range-binned LDS tiles, one
 * wavefront or one lane per step segment): each is launched twice, the fastest
 * is kept for later launches; *mode_out = 1 tiles, 2 wavefront, 3 lane. */
int bhc_problem_autotune(bhc_ctx *ctx, bhc_dev_problem *dp, int *mode_out);
/* Force a scatter mapping (0 = automatic choice, 1 tiles, 2 wavefront, 3 lane). */
int bhc_problem_set_scatter(bhc_dev_problem *dp, int mode);
int bhc_ctx_synchronize(bhc_ctx *ctx);
/* Measurement aid with no counterpart in the reference: one launch of a plain
 * 16-byte-per-lane copy kernel (k_copy16) over `bytes` (a multiple of 16) on
 * the context's stream, *ms its duration. Profiler byte counters read for it
 * give the factor that turns their units into bytes. */
int bhc_copy_calibrate(bhc_ctx *ctx, long long bytes, double *ms);
/* Duration (ms) of the tracer kernel of the last launch on this context,
 * from HIP events recorded on the context's stream around it. */
int bhc_last_kernel_ms(bhc_ctx *ctx, double *ms);
/* Split of the last launch: trace kernel of the first pass, scatter kernel of
 * the first pass (ms), and the number of trace/scatter passes it took. */
int bhc_last_phase_ms(bhc_ctx *ctx, double *trace_ms, double *scatter_ms, int *passes);

/* Launch tuning (rays per workgroup, workgroups per CU); 0 = autotuned default. */
int bhc_ctx_set_launch(bhc_ctx *ctx, int threads_per_block, int blocks_per_cu);

/* ---- progress and non-blocking runs -------------------------------------------------
 * Progress of the context's current or last run: rays finished so far and the
 * run's ray count (bhcInternal::completedRayCount / totalJobs,
 * src/common.hpp:438-440, read by get_percent_progress, include/bhc/bhc.hpp:345-347
 * and src/api.cpp:780-790). May be called from any thread while a run is in
 * flight; done never decreases during a run and equals total once it ends.
 * Trace kernels report finished rays in steps of 256; ray-mode runs report at
 * their end. */
int bhc_progress(bhc_ctx *ctx, int64_t *done, int64_t *total);

/* Non-blocking field runs (extsetup_blocking(params, false); the reference
 * calls completedCallback when the last worker ends,
 * src/mode/fieldimpl.cpp.in:46-52,66-71): the run is performed by a worker
 * thread of the context and returns at once (0, or -1 if it could not start);
 * `completed(user)` is called on the worker thread after the outputs have
 * been written. `prob` and `out` must stay valid until then. While the run is
 * in flight every other call that uses the context fails with an error
 * (bhc_progress, bhc_ctx_busy and bhc_ctx_wait excepted); the callback must not
 * start runs or wait on its own context. bhc_ctx_wait joins the worker and
 * returns the run's status (its message by bhc_last_error). The 3D variant,
 * bhc_field_run3d_async, is declared with the 3D entry points. */
typedef void (*bhc_completed_fn)(void *user);
int bhc_field_run2d_async(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_field_out *out, bhc_completed_fn completed,
                          void *user);
int bhc_ctx_busy(bhc_ctx *ctx);
int bhc_ctx_wait(bhc_ctx *ctx);

/* ---- batched sweeps ---------------------------------------------------------------
 * Many independent TL problems (run types C / S / I, one precision) traced in
 * one persistent launch: rays of all problems are taken from device-side work
 * counters, each problem's field stays in HBM until downloaded. Replaces a
 * caller's loop of RunFieldModesImpl calls over environments
 * (src/mode/fieldimpl.hpp:27-28), one per problem. */
typedef struct bhc_sweep bhc_sweep;
/* Upload `nprob` preprocessed problems; with_stats keeps per-ray step / bounce
 * counts and end points. */
int bhc_sweep_upload2d(bhc_ctx *ctx, const bhc_problem2d *const *probs, int nprob, int with_stats,
                       bhc_sweep **out);
/* Zero every field and trace + scatter all problems (blocks until done). */
int bhc_sweep_launch(bhc_ctx *ctx, bhc_sweep *sw);
/* As bhc_sweep_launch for problems [p0, p1) only: their fields and counters
 * are zeroed and traced in one persistent launch; the other problems' fields
 * are left as they are. Error if the range is empty or outside the sweep. */
int bhc_sweep_launch_range(bhc_ctx *ctx, bhc_sweep *sw, int p0, int p1);
/* Field, errors / warnings, step count and (with_stats) per-ray statistics of
 * problem iprob; out->field holds NSz * NRz_per_range * NRr complex floats. */
int bhc_sweep_download(bhc_ctx *ctx, bhc_sweep *sw, int iprob, bhc_field_out *out);
/* Kernel times (HIP events, summed over passes), passes, total steps and
 * receiver contributions of the last launch. With graph replay (the default)
 * trace_ms holds the whole pass and scatter_ms is 0. */
int bhc_sweep_stats(bhc_sweep *sw, double *trace_ms, double *scatter_ms, int *passes, long long *steps,
                    long long *contrib);
/* Launch-configuration search on the sweep's first 64 problems: scatter grid,
 * scatter mapping (one wavefront per segment, one lane per segment, or 64
 * segments' receivers flattened over a wavefront) and LDS staging of
 * the problems' tables; the fastest is kept for later launches and reported. */
int bhc_sweep_autotune(bhc_ctx *ctx, bhc_sweep *sw, int *scatter_blocks, int *lds_bytes, int *scatter_lanes);
/* 1 (default): each pass (counter reset, trace, scatter) is replayed from a
 * captured HIP graph; 0: enqueued kernel by kernel with per-kernel timing. */
int bhc_sweep_set_graph(bhc_sweep *sw, int use_graph);
/* Scatter mapping of the sweep's passes: 0 one wavefront per segment, 1 one
 * lane per segment, 2 the receivers of 64 segments flattened over a wavefront,
 * 3 fused: the trace kernel evaluates each step's receivers itself (no
 * segment buffer, no scatter kernel, one pass) (bhc_sweep_autotune picks
 * among them). */
int bhc_sweep_set_scatter(bhc_sweep *sw, int mapping);
/* Scatter grid (workgroups) and per-workgroup LDS staging budget of the trace
 * (bytes, a multiple of the real size, at most the upload's budget) without a
 * search: e.g. a configuration bhc_sweep_autotune chose before. */
int bhc_sweep_set_launch(bhc_sweep *sw, int scatter_blocks, int lds_bytes);
void bhc_sweep_release(bhc_sweep *sw);

/* ---- 3D (BELLHOP3D geometry: x, y, depth) ------------------------------------------
 * Full 3D ray tracing (R3D = O3D = true): 3D rays with 2x2 p / q matrices and
 * torsion, triangulated bathymetry / altimetry grids, depth-only or hexahedral
 * sound speed, geometric hat / Gaussian beams in Cartesian coordinates, TL. */

/* Altimetry / bathymetry grid (BdryInfoTopBot<true>, src/module/boundary.hpp:107-197):
 * nx x ny nodes, each rectangle split into two triangles along its diagonal.
 * nx == 0 selects the flat default boundary at the halfspace depth
 * (src/module/boundary.hpp:56-78). type[0]: 'R' flat triangles or 'C'
 * curvilinear (node normals interpolated over the cell, cell curvatures at
 * reflections); type[1]: 'S' one set of geoacoustics or 'L' provinces (each
 * cell takes the cP, cS, rho of its lower corner's province, src/boundary.hpp:134-140). */
typedef struct bhc_bdry3d_in {
    char type[2];
    char pad[6];
    int32_t nx, ny;
    const double *x_km;     /* [nx] */
    const double *y_km;     /* [ny] */
    const double *z;        /* [nx][ny] (m), index ix * ny + iy */
    const int32_t *province; /* 'L': [nx][ny] province numbers (1-based), index ix * ny + iy */
    int32_t nprov, pad2;
    const double *prov;     /* 'L': [nprov][5] alphaR, betaR, rho, alphaI, betaI (attenuation in dB/wavelength) */
} bhc_bdry3d_in;

typedef struct bhc_env3d {
    int32_t precision;      /* 64 or 32 */
    int32_t pad0;
    double freq0;
    /* sound speed: 'C', 'N', 'S', 'P' depth-only profiles, or 'H' hexahedral grid */
    char ssp_type;
    char atten_unit[2];
    char pad1[5];
    int32_t ssp_npts, pad2;
    const double *ssp_z, *ssp_alphaR, *ssp_alphaI, *ssp_rho, *ssp_betaR, *ssp_betaI;
    /* 'H' only: the .ssp grid (src/module/ssp.hpp:144-186), x / y in km, c index (ix * ny + iy) * nz + iz */
    int32_t hex_nx, hex_ny, hex_nz, pad3;
    const double *hex_x_km, *hex_y_km, *hex_z, *hex_c;
    double fg_T, fg_S, fg_pH, fg_z_bar;
    bhc_halfspace_in top, bot;
    bhc_bdry3d_in ati, bty;
    bhc_refl_in trc, brc;
    /* sources and receivers (Position): Sx, Sy, Rr in km; receiver bearings in
     * degrees with the last of a full 360-degree sweep already removed */
    int32_t NSx, NSy, NSz, NRz, NRr, Ntheta;
    const double *Sx_km, *Sy_km;
    const float *Sz, *Rz;
    const double *Rr_km;
    const double *theta_deg;
    /* launch declinations alpha and bearings beta (degrees, sorted) */
    int32_t Nalpha, iSingle_alpha, Nbeta, iSingle_beta;
    const double *alpha_deg, *beta_deg;
    char RunType[7];        /* [0] C / S / I; [1] G / ^ / B (Cartesian beams); [5] '3' */
    char BeamType[3];
    char pad4[6];
    double deltas;          /* m, 0 = automatic */
    double box_x_km, box_y_km, box_z;
    int32_t NSBPPts, pad5;
    const double *sbp;
    int32_t NBioLayers, pad6;   /* as in bhc_env2d */
    const double *bio;
} bhc_env3d;

/* Boundary grid node: position and the unit normals of the two triangles of
 * the cell whose lower corner it is (BdryPtFull<true>::x, n1, n2); the unit
 * node normal Noden; for curvilinear grids the cell's curvatures z_xx, z_xy,
 * z_yy and kappa_xx, kappa_xy, kappa_yy (src/module/boundary.hpp:650-830);
 * for province grids the cell's province (0-based; -1 otherwise). */
typedef struct bhc_bdry_node3d {
    double x[3], n1[3], n2[3];
    double Noden[3];
    double z_xx, z_xy, z_yy, kappa_xx, kappa_xy, kappa_yy;
    int32_t province, pad;
} bhc_bdry_node3d;

typedef struct bhc_problem3d {
    int32_t precision;
    char RunType[7];
    char BeamType[4];
    char ssp_type;
    int32_t pad0;
    double freq0;
    /* depth-only profile (the 2D layout of bhc_problem2d) */
    int32_t ssp_npts, pad1;
    const double *ssp_z, *ssp_c, *ssp_cz, *ssp_n2, *ssp_n2z, *ssp_coef, *ssp_rho;
    /* hexahedral grid, meters; c and cz index (ix * ny + iy) * nz + iz and (ix * ny + iy) * (nz - 1) + iz */
    int32_t hex_nx, hex_ny, hex_nz, pad2;
    const double *hex_x, *hex_y, *hex_z, *hex_c, *hex_cz;
    bhc_hs2d top_hs, bot_hs;
    int32_t ati_nx, ati_ny, bty_nx, bty_ny;
    const bhc_bdry_node3d *ati, *bty;
    char ati_type[2], bty_type[2];  /* [0] 'R' / 'C', [1] 'S' / 'L' */
    int32_t ati_nprov, bty_nprov;
    const double *ati_prov, *bty_prov; /* [nprov][5] cP (re, im), cS (re, im), rho of each province */
    int32_t trc_npts, brc_npts;
    const double *trc, *brc;
    int32_t NSx, NSy, NSz, NRz, NRr, Ntheta;
    const double *Sx, *Sy;      /* m */
    const float *Sz, *Rz;
    const double *Rr;           /* m */
    const double *t_rcvr;       /* [Ntheta][2] cos, sin of the bearings (host libm) */
    double Delta_r;
    int32_t Nalpha, iSingle_alpha, Nbeta, iSingle_beta;
    const double *alpha, *alpha_cos, *alpha_sin;   /* rad; cos / sin from host libm */
    const double *beta, *beta_cos, *beta_sin;
    double Dalpha, Dbeta;
    double deltas, box_x, box_y, box_z;
    int32_t NSBPPts, pad3;
    const double *sbp;
    const double *theta;        /* [Ntheta] receiver bearings, deg (Position::theta) */
} bhc_problem3d;

typedef struct bhc_problem3d_owner bhc_problem3d_owner;
/* 3D preprocessing: the Preprocess() steps of src/module/{ssp.hpp:453-489,
 * boundary.hpp:499-527 + 600-705, beaminfo.hpp:221-238, rayangles.hpp:157-190,
 * rcvrbearings.hpp:98-110, sxsy.hpp:105-112, rcvrranges.hpp:74-85}. */
int bhc_preprocess3d(const bhc_env3d *env, bhc_problem3d_owner **out);
const bhc_problem3d *bhc_problem3d_view(const bhc_problem3d_owner *p);
void bhc_problem3d_free(bhc_problem3d_owner *p);
const char *bhc_host3d_last_error(void);

/* 3D field run, RunFieldModesImpl<CFG, true, true> (src/mode/fieldimpl.hpp:27-28),
 * with Cartesian ('G', 'B') or ray-centred ('g', 'b', Step_InfluenceGeoRayCen
 * src/influence.hpp:1030-1194; equally spaced receiver ranges) geometric beams,
 * for TL (C / S / I), arrivals (A / a: out->arr and out->narr over the receivers
 * in field order, AddArr src/arrivals.hpp:45-130 with azimuths) and eigen hits
 * (E: RecordEigenHit src/eigenrays.hpp:24-41 with itheta, isx, isy, ibeta):
 * out->field is [NSz][NSx][NSy][Ntheta][NRz][NRr] complex floats, caller-zeroed
 * (GetFieldAddr, src/common.hpp:393-405); per-ray statistics as in 2D, with
 * ray_xend holding [rays][3] end points. Rays are numbered like GetJobIndices
 * (src/common.hpp:365-391): alpha fastest, then beta, Sy, Sx, Sz. */
int bhc_field_run3d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out);
/* Non-blocking bhc_field_run3d, with the semantics of bhc_field_run2d_async. */
int bhc_field_run3d_async(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out, bhc_completed_fn completed,
                          void *user);
/* 3D ray run type (RunRayMode, src/mode/ray.cpp:167-188, with MainRayMode
 * src/trace.hpp:482-520 in 3D): every ray of the fan in job order, up to
 * max_points_per_ray points each. */
typedef struct bhc_ray_out3d {
    bhc_ray_pt3d *points;   /* [NRays][max_points_per_ray] */
    int32_t *nsteps;        /* [NRays] */
    int32_t max_points_per_ray;
    int32_t pad;
    uint32_t err, warn, err_count, warn_count;
} bhc_ray_out3d;
int bhc_ray_run3d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_ray_out3d *out);
/* 3D eigenray re-trace (EigenModePostWorker, src/mode/eigen.cpp:24-48): ray k
 * starts from hits[k]'s source and launch angles and stops after step
 * hits[k].is (MainRayMode's Nsteps limit, src/trace.hpp:510-513). */
int bhc_ray_run3d_list(bhc_ctx *ctx, const bhc_problem3d *prob, const bhc_eigen_hit *hits, int32_t nhits,
                       bhc_ray_out3d *out);
/* .ray file of a 3D ray run (OpenRAYFile + WriteRay, src/mode/ray.hpp:168-244):
 * nrays rays in job order. */
int bhc_write_ray3d(const char *path, const bhc_problem3d *prob, const char *title, const bhc_ray_pt3d *pts,
                    const int32_t *nsteps, int32_t max_points, int32_t nrays);
/* ScalePressure for 3D geometric beams (src/influence.hpp:1556-1627): incoherent
 * and semicoherent fields become sqrt(intensity); coherent fields are unchanged.
 * An Nx2D problem (RunType[5] '2', R3D = false) then takes the 2D cylindrical
 * spreading -1 / sqrt(|r|) of each receiver range on every bearing plane. */
int bhc_scale_pressure3d(const bhc_problem3d *prob, float *field);

/* ---- Nx2D: 2D rays on the bearing planes of a 3D environment --------------- */
/* The reference's Nx2D mode (O3D = true, R3D = false; bellhopcxx -Nx2D,
 * run<true, false> of include/bhc/bhc.hpp:308-316). The problem comes from
 * bhc_preprocess3d of an environment with RunType[5] == '2'; for every run
 * type but 'R' its launch bearings are then the receiver bearings
 * (src/module/rayangles.hpp:163-178), so that ray ibeta lies on bearing plane
 * itheta = ibeta (src/influence.hpp:439). Rays are 2D (range along the
 * bearing, depth) and see the 3D boundaries and sound speed along their plane
 * (Origin / RayToOceanX / OceanToRayX, src/common_run.hpp:221-317; Step
 * src/step.hpp:716-765; SSP projection src/ssp.hpp:449-462; Reflect
 * src/reflect.hpp:94-140). Geometric beams (G / B Cartesian, g / b
 * ray-centred), rectilinear boundary grids.
 *
 * Field run (RunFieldModesImpl<CFG, true, false>, src/mode/fieldimpl.hpp:27-28):
 * out->field is [NSz][NSx][NSy][Ntheta][NRz][NRr] complex floats, caller-zeroed
 * (GetFieldAddr, src/common.hpp:393-405); arrivals carry SrcAzimAngle =
 * RcvrAzimAngle = the bearing; eigen hits itheta = ibeta with isx, isy, isz;
 * ray_xend holds [rays][3] ocean end points. Rays are numbered as in 3D. */
int bhc_field_runnx2d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out);
/* Non-blocking bhc_field_runnx2d, with the semantics of bhc_field_run2d_async. */
int bhc_field_runnx2d_async(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out, bhc_completed_fn completed,
                            void *user);
/* Nx2D ray run type (RunRayMode, src/mode/ray.cpp:167-188, MainRayMode
 * src/trace.hpp:482-520 with O3D && !R3D): every ray of the beta fan in job
 * order, points in ray coordinates (rayPt<false>); origin[4 * ray + 0..3] =
 * source x, y (m) and cos, sin of the ray's bearing, which map a point to the
 * ocean as (x + r cos, y + r sin, z) (RayToOceanX; the .ray writer's form). */
int bhc_ray_runnx2d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_ray_out *out, double *origin);
/* PostProcessArrivals for 3D (src/mode/arr.cpp:24-73 with R3D): counts clamped
 * to max_narr, amplitudes unscaled, maximum count per source (isz, isx, isy)
 * written to maxn[NSz * NSx * NSy]. */
int bhc_postprocess_arrivals3d(const bhc_problem3d *prob, bhc_arrival *arr, int32_t *narr, int32_t max_narr,
                               int32_t *maxn);
/* 3D .arr file (WriteOutArrivals<true>, src/mode/arr.cpp:88-238): sources x / y
 * / z, receivers z / r / bearings, then per source and per (bearing, depth,
 * range) the arrivals with source and receiver azimuths; ASCII (binary = 0,
 * run type 'A') or Fortran unformatted ('a'). */
int bhc_write_arr3d(const char *path, const bhc_problem3d *prob, const bhc_arrival *arr, const int32_t *narr,
                    int32_t max_narr, const int32_t *maxn, int binary);

/* ---- post-processing and file layouts (host) ----------------------------------- */

/* ScalePressure + PostProcessTL for 2D (src/influence.hpp:1556-1627, src/mode/tl.cpp:106-180). */
int bhc_scale_pressure2d(const bhc_problem2d *prob, float *field);
/* .shd file (WriteHeader + WriteOutTL, src/mode/tl.cpp:38-101, 211-248). */
int bhc_write_shd2d(const char *path, const bhc_problem2d *prob, const char *title, const float *field);
const char *bhc_host_last_error(void);
/* PostProcessArrivals for 2D (src/mode/arr.cpp:24-73), in place: counts
 * clamped to max_narr, amplitudes scaled by 1/sqrt(r) (1e5 at r = 0, 4 sqrt(pi)
 * for a line source), maximum count per source written to maxn[NSz]. */
int bhc_postprocess_arrivals2d(const bhc_problem2d *prob, bhc_arrival *arr, int32_t *narr, int32_t max_narr,
                               int32_t *maxn);
/* .arr file (WriteOutArrivals, src/mode/arr.cpp:88-238): list-directed ASCII
 * (run types 'A' / 'V', binary = 0) or Fortran unformatted ('a' / 'v'). */
int bhc_write_arr2d(const char *path, const bhc_problem2d *prob, const bhc_arrival *arr, const int32_t *narr,
                    int32_t max_narr, const int32_t *maxn, int binary);
/* .ray file (OpenRAYFile + WriteRay, src/mode/ray.hpp:168-244): nrays rays of
 * max_points points each (the first nsteps[r] valid) and their launch angles
 * in degrees. */
int bhc_write_ray2d(const char *path, const bhc_problem2d *prob, const char *title, const bhc_ray_pt2d *pts,
                    const int32_t *nsteps, int32_t max_points, const double *decl_deg, int32_t nrays);
/* The list-directed writer of the .arr / .ray files as a stream (LDOFile,
 * src/util/ldio.hpp:409-584, FORTRAN_OUTPUT style). */
typedef struct bhc_ldfile bhc_ldfile;
bhc_ldfile *bhc_ld_open(const char *path);
void bhc_ld_i32(bhc_ldfile *f, int32_t v);
void bhc_ld_f32(bhc_ldfile *f, float v);
void bhc_ld_f64(bhc_ldfile *f, double v);
void bhc_ld_str(bhc_ldfile *f, const char *s);
void bhc_ld_char(bhc_ldfile *f, char c);
int bhc_ld_close(bhc_ldfile *f);

/* ---- TL surrogate (no counterpart in the reference) ------------------------------
 * An ensemble of MLPs (16 normalised inputs -> 64 features -> 6 x 256 GELU ->
 * n_out outputs) trained on tracer fields, forward and backward on the matrix
 * cores in BF16 with FP32 accumulation, Adam on FP32 master weights. Feature
 * front 0: sin / cos of 32 projections of the inputs; front 1: the 16 inputs
 * and sin / cos of the first 24 projections. n_out = 1 (a coordinate network:
 * TL at one receiver) or a multiple of 32 up to 256 (the coefficients of a
 * field in a fixed basis). */
typedef struct bhc_mlp bhc_mlp;
int bhc_mlp_create_ex(int device, int members, int max_batch, int n_out, int front, unsigned long long seed,
                      float sigma, bhc_mlp **out);
int bhc_mlp_create(int device, int members, int max_batch, unsigned long long seed, float sigma,
                   bhc_mlp **out);                              /* n_out 1, front 0 */
void bhc_mlp_destroy(bhc_mlp *m);
const char *bhc_mlp_last_error(void);
int bhc_mlp_num_params(void);                                   /* per member, n_out 1 */
long long bhc_mlp_param_count(bhc_mlp *m);                      /* per member */
int bhc_mlp_get_params(bhc_mlp *m, float *params, float *fourier); /* [members][P], [32][16] */
int bhc_mlp_set_params(bhc_mlp *m, const float *params);
int bhc_mlp_set_fourier(bhc_mlp *m, const float *B);            /* [32][16] */
int bhc_mlp_predict(bhc_mlp *m, const float *x, long long n, float *out); /* x [n][16] -> out [members][n][n_out] */
int bhc_mlp_gradients(bhc_mlp *m, const float *x, const float *y, int n, float *grads, float *loss); /* y [n][n_out] */
int bhc_mlp_set_dataset(bhc_mlp *m, const float *X, const float *Y, long long n);      /* Y [n][n_out] */
int bhc_mlp_train(bhc_mlp *m, int steps, int batch, float lr, unsigned long long seed, float *loss_out);
double bhc_mlp_last_ms(bhc_mlp *m);
double bhc_mlp_step_flops(bhc_mlp *m, int batch);

/* ---- measurement (no counterpart in the reference) ------------------------------ */

/* Algorithmic floating-point operation counts of the TL path for `prob`: the
 * same tracer core instantiated on the host with a counting scalar type, run
 * over `nsample` rays spread over the fan. out[0..5] = operations per ray
 * (initialisation), per step (step + influence prelude), per accepted receiver
 * contribution, and the sample's rays, steps and contributions. */
int bhc_opcount2d(const bhc_problem2d *prob, int nsample, double *out);
/* The same for the 3D path (per ray, per step, per accepted contribution). */
int bhc_opcount3d(const bhc_problem3d *prob, int nsample, double *out);

#ifdef __cplusplus
}
#endif
#endif /* BHC_HIP_H */
