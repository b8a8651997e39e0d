// trace2d.hip -- MI355X (gfx950) kernels of the 2D ray/beam tracer and the
// device half of the C ABI (include/bhc_hip.h).
//
// Launch model (replaces FieldModesKernel, src/mode/fieldimpl.cu.in:29-64):
//  * persistent workgroups; every lane owns one ray for the ray's whole life
//    and pulls the next ray index from a device-side work counter as soon as
//    its ray terminates (one atomic per wavefront refill), so the lanes of a
//    wavefront stay busy although rays finish after very different numbers of
//    steps;
//  * the SSP knots / coefficients and the boundary nodes of the problem are
//    staged in LDS once per workgroup;
//  * after each step, every lane reduces its step segment to a descriptor
//    (SegDesc) and the wavefront scatters the segments' receiver work items
//    cooperatively: the descriptor of lane j is broadcast with readlane into
//    scalar registers and the 64 lanes split its receivers. The receiver
//    loops, not the integrator, dominate dense TL grids.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>

#include "../../include/bhc_hip.h"
#include "core2d.h"
#include "pack2d.h"

using namespace bhc2d;

namespace {

thread_local std::string g_last_error;

int fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return -1;
}

#define HIPCHECK(x) \
    do { \
        hipError_t _e = (x); \
        if(_e != hipSuccess) return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while(0)

constexpr int MAXN = 100000;   // include/bhc/structs.hpp:66
constexpr int WAVE = 64;

// Arrival candidate: one raw AddArr call, ordered by (ray job, step).
struct ArrCand {
    long long key;
    double a, phase, dre, dim; // full precision: AddArr compares these before rounding
    float src, rcv;
    int32_t ntop, nbot;
};

struct DevOut {
    float *field;
    ArrCand *cand;
    int32_t *ncand;
    int32_t cand_cap;
    bhc_eigen_hit *hits;
    int32_t *neigen;
    int32_t eigen_memsize;
    int32_t *stats;             // [njobs][3] nsteps, ntop, nbot (or null)
    double *xend;               // [njobs][2]
    bhc_ray_pt2d *points;       // ray mode
    int32_t max_points;
    const int32_t *job_alpha;   // ray mode: explicit ray list (or null)
    const int32_t *job_sz;
    const int32_t *nsteps_in;   // eigen re-trace limits (or null)
    int32_t *nsteps_out;
    unsigned long long *counters; // [0] steps, [1] contributions
    uint32_t *errw;             // err, warn, nerr, nwarn
    int32_t *work;
    int32_t njobs;
};

template<typename real> struct DevSink {
    const DevOut *O;
    const Prob2D<real> *P;
    unsigned long long nc;
    __device__ void add_field(size_t idx, float re, float im)
    {
        unsafeAtomicAdd(O->field + 2 * idx, re);
        unsafeAtomicAdd(O->field + 2 * idx + 1, im);
    }
    __device__ void add_arrival(size_t idx, real amp, real phase, cpx<real> delay, real src, real rcv, int ntop,
                                int nbot, int ialpha, int isz, int is)
    {
        int slot = atomicAdd(O->ncand + idx, 1);
        if(slot >= O->cand_cap) return;
        int job = (P->iSingle >= 1) ? isz : isz * P->Nalpha + ialpha;
        ArrCand c;
        c.key = ((long long)job << 17) | (long long)is;
        c.a = (double)amp;
        c.phase = (double)phase;
        c.dre = (double)delay.re;
        c.dim = (double)delay.im;
        c.src = (float)src;
        c.rcv = (float)rcv;
        c.ntop = ntop;
        c.nbot = nbot;
        O->cand[idx * (size_t)O->cand_cap + slot] = c;
    }
    __device__ void add_eigen(int is, int ir, int iz, int isz, int ialpha)
    {
        int mi = atomicAdd(O->neigen, 1);
        if(mi >= O->eigen_memsize) return;
        bhc_eigen_hit h;
        h.is = is;
        h.itheta = 0;
        h.ir = ir;
        h.iz = iz;
        h.isx = 0;
        h.isy = 0;
        h.isz = isz;
        h.ialpha = ialpha;
        h.ibeta = 0;
        O->hits[mi] = h;
    }
    __device__ void count() { ++nc; }
};

// Broadcast lane j's value to the wavefront (scalar registers).
template<typename T> __device__ inline T bcast(T v, int j)
{
    static_assert(sizeof(T) % 4 == 0, "bcast needs 4-byte multiples");
    constexpr int N = sizeof(T) / 4;
    uint32_t w[N];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for(int i = 0; i < N; ++i) w[i] = (uint32_t)__builtin_amdgcn_readlane((int)w[i], j);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

template<typename real, int RT>
__device__ inline void scatter_wave(const SegDesc<real> &d, int lane, const Prob2D<real> &P, const real *G,
                                    DevSink<real> &sink)
{
    unsigned long long m = __ballot(d.nitems > 0);
    while(m) {
        int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        SegDesc<real> u = bcast(d, j);
        for(int k = lane; k < u.nitems; k += WAVE) item<real, RT>(u, k, P, G, P.Rz, sink);
    }
}

template<typename real> __device__ inline const real *stage(const Prob2D<real> &P, unsigned char *lds)
{
    const real *G = P.base;
    if(P.staged_reals <= 0) return G;
    real *L = reinterpret_cast<real *>(lds);
    for(int i = threadIdx.x; i < P.staged_reals; i += blockDim.x) L[i] = G[i];
    __syncthreads();
    return L;
}

__device__ inline void flush_err(const Err &e, const DevOut &O)
{
    if(e.err) atomicOr(O.errw + 0, e.err);
    if(e.warn) atomicOr(O.errw + 1, e.warn);
    if(e.nerr) atomicAdd(O.errw + 2, e.nerr);
    if(e.nwarn) atomicAdd(O.errw + 3, e.nwarn);
}

// Field modes (TL / arrivals / eigen): MainFieldModes, src/trace.hpp:525-589.
template<typename real, int RT>
__global__ void __launch_bounds__(256) k_field2d(Prob2D<real> P, DevOut O)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const real *B = stage(P, lds_raw);
    const real *G = P.base;
    const int lane = threadIdx.x & (WAVE - 1);
    Err e = {0, 0, 0, 0};
    DevSink<real> sink{&O, &P, 0};
    RayPt<real> p0;
    Trace<real> S;
    Infl<real> I;
    int job = 0, isz = 0, ia = 0, is = 0;
    bool active = false, exhausted = false;
    unsigned long long steps = 0;
    for(int iter = 0;; ++iter) {
        if(!exhausted) {
            unsigned long long need = __ballot(!active);
            if(need) {
                int cnt = __popcll(need);
                int leader = __ffsll((unsigned long long)need) - 1;
                int base = 0;
                if(lane == leader) base = atomicAdd(O.work, cnt);
                base = __shfl(base, leader);
                if(base + cnt >= O.njobs) exhausted = true;
                if(!active) {
                    int my = base + __popcll(need & ((1ull << lane) - 1ull));
                    if(my < O.njobs) {
                        job = my;
                        JobIndices(job, P, isz, ia);
                        is = 0;
                        active = RayInit(isz, ia, p0, S, I, P, B, G, e);
                        if(!active && O.stats) {
                            O.stats[3 * job] = 0;
                            O.stats[3 * job + 1] = 0;
                            O.stats[3 * job + 2] = 0;
                            O.xend[2 * job] = 0.0;
                            O.xend[2 * job + 1] = 0.0;
                        }
                    }
                }
            }
        }
        if(!__any(active)) {
            if(exhausted) break;
            continue;
        }
        if((iter & 31) == 31) {
            uint32_t ge = __hip_atomic_load(O.errw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if(ge != 0u || e.err != 0u) {
                exhausted = true;
                active = false;
                continue;
            }
        }
        RayPt<real> p1, p2;
        SegDesc<real> dA, dB;
        dA.nitems = 0;
        dB.nitems = 0;
        bool two = false;
        if(active) {
            two = RayUpdate(p0, p1, p2, S, P, B, G, e);
            prelude(dA, p0, p1, I, is, P, G, isz, ia);
            ++is;
            if(two) {
                prelude(dB, p1, p2, I, is, P, G, isz, ia);
                ++is;
            }
        }
        scatter_wave<real, RT>(dA, lane, P, G, sink);
        scatter_wave<real, RT>(dB, lane, P, G, sink);
        if(active) {
            p0 = two ? p2 : p1;
            int t = RayTerminate(p0, is, S, MAXN, P, e);
            if(t) {
                active = false;
                steps += (unsigned long long)is;
                if(O.stats) {
                    O.stats[3 * job] = is;
                    O.stats[3 * job + 1] = p0.NumTopBnc;
                    O.stats[3 * job + 2] = p0.NumBotBnc;
                    O.xend[2 * job] = (double)p0.x.x;
                    O.xend[2 * job + 1] = (double)p0.x.y;
                }
            }
        }
    }
    flush_err(e, O);
    // wave-reduce the work counters before the global atomics
    for(int off = WAVE / 2; off > 0; off >>= 1) {
        steps += __shfl_xor(steps, off);
        sink.nc += __shfl_xor(sink.nc, off);
    }
    if(lane == 0) {
        atomicAdd(O.counters + 0, steps);
        atomicAdd(O.counters + 1, sink.nc);
    }
}

template<typename real> __device__ inline void store_pt(bhc_ray_pt2d *d, const RayPt<real> &s)
{
    bhc_ray_pt2d r;
    r.p[0] = s.p.x;
    r.p[1] = s.p.y;
    r.q[0] = s.q.x;
    r.q[1] = s.q.y;
    r.NumTopBnc = s.NumTopBnc;
    r.NumBotBnc = s.NumBotBnc;
    r.x[0] = s.x.x;
    r.x[1] = s.x.y;
    r.t[0] = s.t.x;
    r.t[1] = s.t.y;
    r.c = s.c;
    r.Amp = s.Amp;
    r.Phase = s.Phase;
    r.tau[0] = s.tau.re;
    r.tau[1] = s.tau.im;
    *d = r;
}

// Ray run type and eigenray re-trace: MainRayMode, src/trace.hpp:482-520.
template<typename real>
__global__ void __launch_bounds__(256) k_ray2d(Prob2D<real> P, DevOut O)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const real *B = stage(P, lds_raw);
    const real *G = P.base;
    Err e = {0, 0, 0, 0};
    const int M = O.max_points;
    for(int job = blockIdx.x * blockDim.x + threadIdx.x; job < O.njobs; job += gridDim.x * blockDim.x) {
        int isz, ia;
        if(O.job_alpha) {
            ia = O.job_alpha[job];
            isz = O.job_sz[job];
        } else {
            JobIndices(job, P, isz, ia);
        }
        bhc_ray_pt2d *ray = O.points + (size_t)job * M;
        RayPt<real> p0, p1, p2;
        Trace<real> S;
        Infl<real> I;
        int Nsteps = O.nsteps_in ? O.nsteps_in[job] : -1;
        bool ok = RayInit(isz, ia, p0, S, I, P, B, G, e);
        store_pt(ray, p0);
        if(!ok) {
            O.nsteps_out[job] = 1;
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdate(p0, p1, p2, S, P, B, G, e);
            store_pt(ray + is + 1, p1);
            if(two) store_pt(ray + is + 2, p2);
            if(Nsteps >= 0 && is >= Nsteps) {
                Nsteps = is + 2;
                break;
            }
            is += two ? 2 : 1;
            p0 = two ? p2 : p1;
            int t = RayTerminate(p0, is, S, M, P, e);
            if(t == 1) {
                Nsteps = is + 1;
                break;
            }
            if(t == 2) {
                Nsteps = is;
                break;
            }
        }
        O.nsteps_out[job] = Nsteps;
    }
    flush_err(e, O);
}

// Deterministic arrivals: one thread per receiver sorts its candidates by
// (ray, step) -- the order of the reference's single-threaded run -- and
// replays AddArr (src/arrivals.hpp:45-130) on them.
__global__ void __launch_bounds__(256) k_arr_merge(const ArrCand *cand, const int32_t *ncand, int cand_cap,
                                                   bhc_arrival *arr, int32_t *narr, int max_narr, int nrcv,
                                                   int allow_merging, float omega_f, double omega_d, int fp64,
                                                   uint32_t *errw)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if(r >= nrcv) return;
    int n = ncand[r];
    if(n > cand_cap) {
        atomicOr(errw + 1, 1u << BHC_WARN_RAYS_OUTOFMEMORY);
        atomicAdd(errw + 3, 1u);
        n = cand_cap;
    }
    ArrCand *c = const_cast<ArrCand *>(cand) + (size_t)r * cand_cap;
    // insertion sort: candidates arrive nearly sorted (rays are handed out in order)
    for(int i = 1; i < n; ++i) {
        ArrCand v = c[i];
        int j = i - 1;
        while(j >= 0 && c[j].key > v.key) {
            c[j + 1] = c[j];
            --j;
        }
        c[j + 1] = v;
    }
    bhc_arrival *A = arr + (size_t)r * max_narr;
    int Nt = 0;
    for(int i = 0; i < n; ++i) {
        const ArrCand &v = c[i];
        if(!allow_merging) {
            if(Nt >= max_narr) break;
            bhc_arrival &a = A[Nt++];
            a.a = (float)v.a;
            a.Phase = (float)v.phase;
            a.delay[0] = (float)v.dre;
            a.delay[1] = (float)v.dim;
            a.SrcDeclAngle = v.src;
            a.SrcAzimAngle = NAN;
            a.RcvrDeclAngle = v.rcv;
            a.RcvrAzimAngle = NAN;
            a.NTopBnc = v.ntop;
            a.NBotBnc = v.nbot;
            continue;
        }
        // IsSecondStepOfPair, arrivals.hpp:31-39, in the run's real type
        bool second = false;
        if(Nt >= 1) {
            const float PhaseTol = 0.05f;
            const bhc_arrival &b = A[Nt - 1];
            if(fp64) {
                double dr = v.dre - (double)b.delay[0], di = v.dim - (double)b.delay[1];
                second = omega_d * hypot(dr, di) < (double)PhaseTol
                    && fabs((double)b.Phase - v.phase) < (double)PhaseTol;
            } else {
                float dr = (float)v.dre - b.delay[0], di = (float)v.dim - b.delay[1];
                second = omega_f * hypotf(dr, di) < PhaseTol && fabsf(b.Phase - (float)v.phase) < PhaseTol;
            }
        }
        if(!second) {
            int iArr;
            if(Nt >= max_narr) {
                iArr = -1;
                double weakest = v.a;
                for(int k = 0; k < max_narr; ++k)
                    if((double)A[k].a < weakest) {
                        weakest = A[k].a;
                        iArr = k;
                    }
                if(iArr < 0) continue;
            } else {
                iArr = Nt++;
            }
            bhc_arrival &a = A[iArr];
            a.a = (float)v.a;
            a.Phase = (float)v.phase;
            a.delay[0] = (float)v.dre;
            a.delay[1] = (float)v.dim;
            a.SrcDeclAngle = v.src;
            a.SrcAzimAngle = NAN;
            a.RcvrDeclAngle = v.rcv;
            a.RcvrAzimAngle = NAN;
            a.NTopBnc = v.ntop;
            a.NBotBnc = v.nbot;
        } else {
            bhc_arrival &a = A[Nt - 1];
            float amp = (float)v.a;
            float AmpTot = a.a + amp;
            float w1 = a.a / AmpTot, w2 = amp / AmpTot;
            a.delay[0] = w1 * a.delay[0] + w2 * (float)v.dre;
            a.delay[1] = w1 * a.delay[1] + w2 * (float)v.dim;
            a.a = AmpTot;
            a.SrcDeclAngle = w1 * a.SrcDeclAngle + w2 * v.src;
            a.SrcAzimAngle = w1 * a.SrcAzimAngle + w2 * (float)NAN;
            a.RcvrDeclAngle = w1 * a.RcvrDeclAngle + w2 * v.rcv;
            a.RcvrAzimAngle = w1 * a.RcvrAzimAngle + w2 * (float)NAN;
        }
    }
    narr[r] = Nt;
}

} // namespace

// ===========================================================================
// Host side of the C ABI
// ===========================================================================
struct bhc_ctx {
    int device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    int num_cus;
    int lds_max;
    int tpb, bpc;
    char name[256];
};

struct bhc_dev_problem {
    bhc_ctx *ctx;
    int precision;
    char run0;
    int njobs, NSz, NRz_per_range, NRr;
    size_t nrcv;
    void *d_blob;
    float *d_Rz, *d_Sz;
    Prob2D<double> P64;
    Prob2D<float> P32;
    size_t lds_bytes;
    float *d_field;
    ArrCand *d_cand;
    int32_t *d_ncand;
    int cand_cap;
    bhc_arrival *d_arr;
    int32_t *d_narr;
    int max_narr, allow_merging;
    bhc_eigen_hit *d_hits;
    int32_t *d_neigen;
    int eigen_memsize;
    int32_t *d_stats;
    double *d_xend;
    unsigned long long *d_counters;
    uint32_t *d_errw;
    int32_t *d_work;
    double freq0;
};

extern "C" const char *bhc_last_error(void) { return g_last_error.c_str(); }

extern "C" int bhc_device_info(int device, char *name, int name_len, int *num_cus, int *lds_bytes)
{
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    if(name && name_len > 0) snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if(num_cus) *num_cus = prop.multiProcessorCount;
    if(lds_bytes) *lds_bytes = (int)prop.sharedMemPerBlock;
    return 0;
}

extern "C" int bhc_ctx_create(int device, bhc_ctx **out)
{
    *out = nullptr;
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if(device < 0 || device >= n) return fail("bhc_ctx_create: device %d not present (%d devices)", device, n);
    HIPCHECK(hipSetDevice(device));
    bhc_ctx *c = new bhc_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    c->lds_max = (int)prop.sharedMemPerBlock;
    snprintf(c->name, sizeof(c->name), "%s", prop.gcnArchName);
    c->tpb = 256;
    c->bpc = 2;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->ev0));
    HIPCHECK(hipEventCreate(&c->ev1));
    *out = c;
    return 0;
}

extern "C" void bhc_ctx_destroy(bhc_ctx *ctx)
{
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int bhc_ctx_set_launch(bhc_ctx *ctx, int threads_per_block, int blocks_per_cu)
{
    if(threads_per_block != 0 && threads_per_block != 256)
        return fail("bhc_ctx_set_launch: kernels are built for 256 threads per workgroup");
    if(blocks_per_cu < 0 || blocks_per_cu > 64) return fail("bhc_ctx_set_launch: blocks_per_cu out of range");
    ctx->bpc = blocks_per_cu ? blocks_per_cu : 2;
    return 0;
}

extern "C" int bhc_ctx_synchronize(bhc_ctx *ctx)
{
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int bhc_last_kernel_ms(bhc_ctx *ctx, double *ms)
{
    float f = 0.f;
    HIPCHECK(hipEventSynchronize(ctx->ev1));
    HIPCHECK(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = f;
    return 0;
}

namespace {

template<typename T> int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if(n == 0) return 0;
    HIPCHECK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}

int check_problem(const bhc_problem2d *p)
{
    if(!p) return fail("null problem");
    if(p->precision != 64 && p->precision != 32) return fail("precision must be 32 or 64");
    if(p->NSz < 1 || p->NRz < 1 || p->NRr < 1 || p->Nalpha < 1) return fail("empty source/receiver/angle set");
    if(p->NRz_per_range != p->NRz && p->NRz_per_range != 1) return fail("bad NRz_per_range");
    if(p->RunType[4] == 'I' && p->NRz != p->NRr) return fail("irregular grid needs NRz == NRr");
    if(p->iSingle_alpha < 0 || p->iSingle_alpha > p->Nalpha) return fail("iSingle out of range");
    if(p->NSBPPts < 2) return fail("source beam pattern needs 2 points");
    if((long long)p->NSz * p->Nalpha >= (1ll << 31) / 2) return fail("too many rays");
    if(!(p->deltas > 0)) return fail("deltas must be positive after preprocessing");
    return 0;
}

} // namespace

extern "C" void bhc_problem_release(bhc_dev_problem *dp)
{
    if(!dp) return;
    (void)hipSetDevice(dp->ctx->device);
    (void)hipStreamSynchronize(dp->ctx->stream);
    void *ptrs[] = {dp->d_blob, dp->d_Rz,   dp->d_Sz,     dp->d_field,    dp->d_cand, dp->d_ncand,
                    dp->d_arr,  dp->d_narr, dp->d_hits,   dp->d_neigen,   dp->d_stats, dp->d_xend,
                    dp->d_counters, dp->d_errw, dp->d_work};
    for(void *p : ptrs)
        if(p) (void)hipFree(p);
    delete dp;
}

extern "C" int bhc_problem_upload2d(bhc_ctx *ctx, const bhc_problem2d *prob, const bhc_field_out *spec,
                                    bhc_dev_problem **out)
{
    *out = nullptr;
    if(check_problem(prob)) return -1;
    HIPCHECK(hipSetDevice(ctx->device));
    bhc_dev_problem *dp = new bhc_dev_problem();
    memset(dp, 0, sizeof(*dp));
    dp->ctx = ctx;
    dp->precision = prob->precision;
    dp->run0 = prob->RunType[0];
    dp->NSz = prob->NSz;
    dp->NRz_per_range = prob->NRz_per_range;
    dp->NRr = prob->NRr;
    dp->njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * prob->NSz;
    dp->nrcv = (size_t)prob->NSz * prob->NRz_per_range * prob->NRr;
    dp->freq0 = prob->freq0;
    int rc = 0;
    size_t blob_bytes = 0, staged = 0;
    if(prob->precision == 64) {
        std::vector<double> blob;
        std::string err;
        rc = pack<double>(prob, blob, dp->P64, err);
        if(rc) fail("%s", err.c_str());
        if(!rc) {
            blob_bytes = blob.size() * sizeof(double);
            staged = (size_t)dp->P64.staged_reals * sizeof(double);
            rc = dalloc((unsigned char **)&dp->d_blob, blob_bytes);
            if(!rc && hipMemcpy(dp->d_blob, blob.data(), blob_bytes, hipMemcpyHostToDevice) != hipSuccess)
                rc = fail("blob upload failed");
        }
    } else {
        std::vector<float> blob;
        std::string err;
        rc = pack<float>(prob, blob, dp->P32, err);
        if(rc) fail("%s", err.c_str());
        if(!rc) {
            blob_bytes = blob.size() * sizeof(float);
            staged = (size_t)dp->P32.staged_reals * sizeof(float);
            rc = dalloc((unsigned char **)&dp->d_blob, blob_bytes);
            if(!rc && hipMemcpy(dp->d_blob, blob.data(), blob_bytes, hipMemcpyHostToDevice) != hipSuccess)
                rc = fail("blob upload failed");
        }
    }
    if(rc) {
        bhc_problem_release(dp);
        return -1;
    }
    // Stage in LDS when the SSP + boundary tables fit in 64 KiB.
    dp->lds_bytes = (staged <= 65536) ? staged : 0;
    if(!dp->lds_bytes) {
        dp->P64.staged_reals = 0;
        dp->P32.staged_reals = 0;
    }
    dp->P64.base = (const double *)dp->d_blob;
    dp->P32.base = (const float *)dp->d_blob;
    if(dalloc(&dp->d_Rz, prob->NRz) || dalloc(&dp->d_Sz, prob->NSz) || dalloc(&dp->d_counters, 2)
       || dalloc(&dp->d_errw, 4) || dalloc(&dp->d_work, 1)) {
        bhc_problem_release(dp);
        return -1;
    }
    HIPCHECK(hipMemcpy(dp->d_Rz, prob->Rz, prob->NRz * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dp->d_Sz, prob->Sz, prob->NSz * sizeof(float), hipMemcpyHostToDevice));
    dp->P64.Rz = dp->P32.Rz = dp->d_Rz;
    dp->P64.Sz = dp->P32.Sz = dp->d_Sz;
    char r0 = dp->run0;
    bool tl = (r0 == 'C' || r0 == 'S' || r0 == 'I');
    bool arr = (r0 == 'A' || r0 == 'a' || r0 == 'V' || r0 == 'v');
    bool eig = (r0 == 'E' || r0 == 'V' || r0 == 'v');
    if(!tl && !arr && !eig && r0 != 'R') {
        bhc_problem_release(dp);
        return fail("unknown run type '%c'", r0);
    }
    if(tl && dalloc(&dp->d_field, 2 * dp->nrcv)) {
        bhc_problem_release(dp);
        return -1;
    }
    if(arr) {
        dp->max_narr = spec ? spec->max_narr : 0;
        dp->allow_merging = spec ? spec->allow_merging : 1;
        if(dp->max_narr < 1) {
            bhc_problem_release(dp);
            return fail("arrivals run needs max_narr >= 1");
        }
        // candidate capacity per receiver from a 16 GiB budget, at least 8 x MaxNArr
        size_t budget = (size_t)16 << 30;
        size_t cap = budget / (dp->nrcv * sizeof(ArrCand));
        cap = std::min<size_t>(cap, 65536);
        cap = std::max<size_t>(cap, 8 * (size_t)dp->max_narr);
        dp->cand_cap = (int)cap;
        if(dalloc(&dp->d_cand, dp->nrcv * cap) || dalloc(&dp->d_ncand, dp->nrcv)
           || dalloc(&dp->d_arr, dp->nrcv * dp->max_narr) || dalloc(&dp->d_narr, dp->nrcv)) {
            bhc_problem_release(dp);
            return -1;
        }
    }
    if(eig) {
        dp->eigen_memsize = spec ? spec->eigen_memsize : 0;
        if(dp->eigen_memsize < 1) {
            bhc_problem_release(dp);
            return fail("eigenray run needs eigen_memsize >= 1");
        }
        if(dalloc(&dp->d_hits, dp->eigen_memsize) || dalloc(&dp->d_neigen, 1)) {
            bhc_problem_release(dp);
            return -1;
        }
    }
    if(spec && spec->ray_nsteps) {
        if(dalloc(&dp->d_stats, 3 * (size_t)dp->njobs) || dalloc(&dp->d_xend, 2 * (size_t)dp->njobs)) {
            bhc_problem_release(dp);
            return -1;
        }
    }
    *out = dp;
    return 0;
}

namespace {
DevOut make_out(bhc_dev_problem *dp)
{
    DevOut O;
    memset(&O, 0, sizeof(O));
    O.field = dp->d_field;
    O.cand = dp->d_cand;
    O.ncand = dp->d_ncand;
    O.cand_cap = dp->cand_cap;
    O.hits = dp->d_hits;
    O.neigen = dp->d_neigen;
    O.eigen_memsize = dp->eigen_memsize;
    O.stats = dp->d_stats;
    O.xend = dp->d_xend;
    O.counters = dp->d_counters;
    O.errw = dp->d_errw;
    O.work = dp->d_work;
    O.njobs = dp->njobs;
    return O;
}

template<typename real, int RT> int launch_field(bhc_dev_problem *dp, const Prob2D<real> &P, const DevOut &O)
{
    bhc_ctx *c = dp->ctx;
    int waves = (dp->njobs + WAVE - 1) / WAVE;
    int blocks = std::min((waves + 3) / 4, c->num_cus * c->bpc);
    blocks = std::max(blocks, 1);
    hipLaunchKernelGGL((k_field2d<real, RT>), dim3(blocks), dim3(256), dp->lds_bytes, c->stream, P, O);
    HIPCHECK(hipGetLastError());
    return 0;
}
} // namespace

extern "C" int bhc_field_launch(bhc_ctx *ctx, bhc_dev_problem *dp)
{
    if(!dp || dp->ctx != ctx) return fail("bhc_field_launch: problem not uploaded on this context");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if(dp->d_field) HIPCHECK(hipMemsetAsync(dp->d_field, 0, 2 * dp->nrcv * sizeof(float), s));
    if(dp->d_ncand) HIPCHECK(hipMemsetAsync(dp->d_ncand, 0, dp->nrcv * sizeof(int32_t), s));
    if(dp->d_narr) HIPCHECK(hipMemsetAsync(dp->d_narr, 0, dp->nrcv * sizeof(int32_t), s));
    if(dp->d_neigen) HIPCHECK(hipMemsetAsync(dp->d_neigen, 0, sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(dp->d_counters, 0, 2 * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(dp->d_errw, 0, 4 * sizeof(uint32_t), s));
    HIPCHECK(hipMemsetAsync(dp->d_work, 0, sizeof(int32_t), s));
    DevOut O = make_out(dp);
    char r0 = dp->run0;
    int rt = (r0 == 'C' || r0 == 'S' || r0 == 'I') ? RT_TL : (r0 == 'E') ? RT_EIG : RT_ARR;
    HIPCHECK(hipEventRecord(ctx->ev0, s));
    int rc;
    if(dp->precision == 64) {
        rc = rt == RT_TL ? launch_field<double, RT_TL>(dp, dp->P64, O)
            : rt == RT_ARR ? launch_field<double, RT_ARR>(dp, dp->P64, O)
                           : launch_field<double, RT_EIG>(dp, dp->P64, O);
    } else {
        rc = rt == RT_TL ? launch_field<float, RT_TL>(dp, dp->P32, O)
            : rt == RT_ARR ? launch_field<float, RT_ARR>(dp, dp->P32, O)
                           : launch_field<float, RT_EIG>(dp, dp->P32, O);
    }
    HIPCHECK(hipEventRecord(ctx->ev1, s));
    if(rc) return rc;
    if(rt == RT_ARR) {
        int nrcv = (int)dp->nrcv;
        double om = 2.0 * M_PI * dp->freq0;
        hipLaunchKernelGGL(k_arr_merge, dim3((nrcv + 255) / 256), dim3(256), 0, s, dp->d_cand, dp->d_ncand,
                           dp->cand_cap, dp->d_arr, dp->d_narr, dp->max_narr, nrcv, dp->allow_merging,
                           (float)(2.0f * (float)M_PI * (float)dp->freq0), om, dp->precision == 64 ? 1 : 0,
                           dp->d_errw);
        HIPCHECK(hipGetLastError());
    }
    return 0;
}

extern "C" int bhc_field_download(bhc_ctx *ctx, bhc_dev_problem *dp, bhc_field_out *out)
{
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if(out->field && dp->d_field)
        HIPCHECK(hipMemcpy(out->field, dp->d_field, 2 * dp->nrcv * sizeof(float), hipMemcpyDeviceToHost));
    if(out->arr && dp->d_arr) {
        if(out->max_narr != dp->max_narr) return fail("download: max_narr differs from the upload spec");
        HIPCHECK(hipMemcpy(out->arr, dp->d_arr, dp->nrcv * dp->max_narr * sizeof(bhc_arrival), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(out->narr, dp->d_narr, dp->nrcv * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    if(out->hits && dp->d_hits) {
        int n = 0;
        HIPCHECK(hipMemcpy(&n, dp->d_neigen, sizeof(int32_t), hipMemcpyDeviceToHost));
        int m = std::min(n, std::min(out->eigen_memsize, dp->eigen_memsize));
        if(m > 0) HIPCHECK(hipMemcpy(out->hits, dp->d_hits, m * sizeof(bhc_eigen_hit), hipMemcpyDeviceToHost));
        if(out->neigen) *out->neigen = n;
    }
    if(out->ray_nsteps && dp->d_stats) {
        std::vector<int32_t> st(3 * (size_t)dp->njobs);
        HIPCHECK(hipMemcpy(st.data(), dp->d_stats, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for(int j = 0; j < dp->njobs; ++j) {
            out->ray_nsteps[j] = st[3 * j];
            if(out->ray_ntop) out->ray_ntop[j] = st[3 * j + 1];
            if(out->ray_nbot) out->ray_nbot[j] = st[3 * j + 2];
        }
        if(out->ray_xend)
            HIPCHECK(hipMemcpy(out->ray_xend, dp->d_xend, 2 * (size_t)dp->njobs * sizeof(double), hipMemcpyDeviceToHost));
    }
    unsigned long long cnt[2];
    uint32_t ew[4];
    HIPCHECK(hipMemcpy(cnt, dp->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(ew, dp->d_errw, sizeof(ew), hipMemcpyDeviceToHost));
    out->total_steps = (int64_t)cnt[0];
    out->total_contrib = (int64_t)cnt[1];
    out->err = ew[0];
    out->warn = ew[1];
    out->err_count = ew[2];
    out->warn_count = ew[3];
    return 0;
}

extern "C" int bhc_field_run2d(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_field_out *out)
{
    bhc_dev_problem *dp = nullptr;
    if(bhc_problem_upload2d(ctx, prob, out, &dp)) return -1;
    int rc = bhc_field_launch(ctx, dp);
    if(!rc) rc = bhc_field_download(ctx, dp, out);
    bhc_problem_release(dp);
    return rc;
}

extern "C" int bhc_ray_run2d(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_ray_out *out)
{
    return bhc_ray_run2d_list(ctx, prob, out, nullptr, nullptr, nullptr, -1);
}

extern "C" int bhc_ray_run2d_list(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_ray_out *out, const int32_t *ialpha,
                                  const int32_t *isz, const int32_t *nsteps_in, int32_t nrays)
{
    if(check_problem(prob)) return -1;
    if(out->max_points_per_ray < 3) return fail("max_points_per_ray must be >= 3");
    HIPCHECK(hipSetDevice(ctx->device));
    bhc_field_out spec;
    memset(&spec, 0, sizeof(spec));
    bhc_problem2d p2 = *prob;
    p2.RunType[0] = 'R';
    bhc_dev_problem *dp = nullptr;
    if(bhc_problem_upload2d(ctx, &p2, &spec, &dp)) return -1;
    int njobs = nrays >= 0 ? nrays : dp->njobs;
    DevOut O = make_out(dp);
    O.njobs = njobs;
    O.max_points = out->max_points_per_ray;
    int32_t *d_ia = nullptr, *d_isz = nullptr, *d_nin = nullptr, *d_nout = nullptr;
    bhc_ray_pt2d *d_pts = nullptr;
    int rc = 0;
    size_t npts = (size_t)njobs * out->max_points_per_ray;
    if(dalloc(&d_pts, npts) || dalloc(&d_nout, njobs)) rc = -1;
    if(!rc && ialpha) {
        if(dalloc(&d_ia, njobs) || dalloc(&d_isz, njobs)) rc = -1;
        else {
            hipMemcpy(d_ia, ialpha, njobs * sizeof(int32_t), hipMemcpyHostToDevice);
            hipMemcpy(d_isz, isz, njobs * sizeof(int32_t), hipMemcpyHostToDevice);
        }
    }
    if(!rc && nsteps_in) {
        if(dalloc(&d_nin, njobs)) rc = -1;
        else hipMemcpy(d_nin, nsteps_in, njobs * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    if(!rc && njobs > 0) {
        O.points = d_pts;
        O.job_alpha = d_ia;
        O.job_sz = d_isz;
        O.nsteps_in = d_nin;
        O.nsteps_out = d_nout;
        hipMemsetAsync(dp->d_errw, 0, 4 * sizeof(uint32_t), ctx->stream);
        hipMemsetAsync(d_pts, 0, npts * sizeof(bhc_ray_pt2d), ctx->stream);
        int blocks = std::max(1, std::min((njobs + 255) / 256, ctx->num_cus * ctx->bpc));
        if(dp->precision == 64)
            hipLaunchKernelGGL(k_ray2d<double>, dim3(blocks), dim3(256), dp->lds_bytes, ctx->stream, dp->P64, O);
        else
            hipLaunchKernelGGL(k_ray2d<float>, dim3(blocks), dim3(256), dp->lds_bytes, ctx->stream, dp->P32, O);
        if(hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail("ray kernel failed");
        } else {
            hipMemcpy(out->points, d_pts, npts * sizeof(bhc_ray_pt2d), hipMemcpyDeviceToHost);
            hipMemcpy(out->nsteps, d_nout, njobs * sizeof(int32_t), hipMemcpyDeviceToHost);
            uint32_t ew[4];
            hipMemcpy(ew, dp->d_errw, sizeof(ew), hipMemcpyDeviceToHost);
            out->err = ew[0];
            out->warn = ew[1];
            out->err_count = ew[2];
            out->warn_count = ew[3];
        }
    }
    void *fr[] = {d_ia, d_isz, d_nin, d_nout, d_pts};
    for(void *p : fr)
        if(p) (void)hipFree(p);
    bhc_problem_release(dp);
    return rc;
}
