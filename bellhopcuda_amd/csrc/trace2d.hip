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
//    (SegDesc: the receiver range / depth window it can reach and the
//    interpolation constants of the influence) and appends it to a segment
//    buffer in HBM (one atomic per wavefront).
// A second phase scatters the segments' receiver work items:
//  * k_scatter2d: one wavefront per segment, descriptor in scalar registers,
//    the 64 lanes split its receivers, FP32 atomics into the field;
//  * tiled TL path: segments are counting-sorted into receiver-range bins
//    (k_bin_count / k_bin_scan / k_bin_fill); k_scatter_tile accumulates one
//    range tile of the field in LDS and flushes it with one atomic per
//    receiver, which keeps the dense grids' atomics out of L2.
// A full segment buffer ends the pass; the host splits the ray range and
// traces the rest in further passes (the field is only written by the
// scatter, so the split leaves the result unchanged).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>
#include <atomic>
#include <mutex>
#include <thread>

#include "../../include/bhc_hip.h"
#include "core2d.h"
#include "pack2d.h"
#include "nx2d.h"
#include "pack3d.h"

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
    float src, rcv;             // source / receiver declination (deg)
    float srcaz, rcvaz;         // source / receiver azimuth (deg; NaN in 2D)
    int32_t ntop, nbot;
};

// Stored step segment: the descriptor plus padding to a 16-byte multiple.
template<typename real> struct alignas(16) Seg {
    SegDesc<real> d;
};

// Segment buffer layout, structure of arrays: 16-byte word k of slot s's
// descriptor at segs + (k seg_cap + s) 16, so that a wavefront's 64 lanes
// store and load consecutive words (coalesced, one 16-byte access per lane
// and word). Words below SEG_NC hold the core (every field a Cartesian TL
// scatter reads); the rest (ray-centred geometry, arrival and eigen fields)
// is stored only where seg_full. A slot takes sizeof(Seg<real>) bytes of the
// buffer in all.
template<typename real> constexpr int SEG_NC = (int)((seg_core_bytes<real>() + 15) / 16);
template<typename real> constexpr int SEG_NW = (int)((sizeof(SegDesc<real>) + 15) / 16);
static_assert(SEG_NW<double> * 16 <= sizeof(Seg<double>) && SEG_NW<float> * 16 <= sizeof(Seg<float>), "slot size");
template<typename real> union SegWords {
    SegDesc<real> d;
    unsigned long long w[2 * SEG_NW<real>];
};


// Finished-ray progress of a launch (completedRayCount / totalJobs,
// src/mode/fieldimpl.cpp.in:43,62-63): wavefronts add their finished rays to
// a device counter; each time the total crosses a multiple of PROG_QUANTUM it
// is also stored (plain system-scope store, no host atomics) to the
// context's pinned host word, which bhc_progress reads while kernels run.
// Stores may land out of order; the reader keeps the largest value seen.
constexpr unsigned long long PROG_QUANTUM = 256ull;
struct Progress {
    unsigned long long *dev;    // device counter (null: not counted)
    unsigned long long *host;   // pinned host mirror
};

__device__ inline void note_done(const Progress &G, bool fin)
{
    const unsigned long long m = __ballot(fin);
    if(!m || !G.dev) return;
    if((int)(threadIdx.x & (WAVE - 1)) == __ffsll(m) - 1) {
        const unsigned long long c = (unsigned long long)__popcll(m);
        const unsigned long long old = atomicAdd(G.dev, c);
        if((old + c) / PROG_QUANTUM != old / PROG_QUANTUM)
            __hip_atomic_store(G.host, old + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

struct DevOut {
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
    int32_t *work;              // ray work counter (relative to job_begin)
    int32_t job_begin, job_end; // ray range of this trace pass
    void *segs;                 // segment buffer: seg_cap descriptor cores, then seg_cap rests (seg_core / seg_rest)
    uint32_t *seg_count;
    uint32_t seg_cap;
    uint32_t *overflow;
    Progress prog;
};

// First 8 bytes of 16-byte word k of slot s.
__device__ inline unsigned long long *seg_word(const DevOut &O, int k, size_t s)
{
    return reinterpret_cast<unsigned long long *>(O.segs) + 2 * ((size_t)k * O.seg_cap + s);
}
template<typename real> __device__ inline int seg_nitems(const DevOut &O, size_t s)
{
    constexpr size_t off = offsetof(SegDesc<real>, nitems);
    return reinterpret_cast<const int32_t *>(seg_word(O, (int)(off / 16), s))[(off % 16) / 4];
}
template<typename real> __device__ inline void clear_seg(const DevOut &O, size_t s)
{
    constexpr size_t off = offsetof(SegDesc<real>, nitems);
    reinterpret_cast<int32_t *>(seg_word(O, (int)(off / 16), s))[(off % 16) / 4] = 0;
}
template<typename real> __device__ inline void load_seg_words(const DevOut &O, size_t s, int k0, int k1, SegWords<real> &u)
{
#pragma unroll
    for(int k = 0; k < SEG_NW<real>; ++k)
        if(k >= k0 && k < k1) {
            const uint4 v = *reinterpret_cast<const uint4 *>(seg_word(O, k, s));
            u.w[2 * k] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
            u.w[2 * k + 1] = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
        }
}
template<typename real> __device__ inline void load_seg_core(const DevOut &O, size_t s, SegDesc<real> &d)
{
    SegWords<real> u;
    load_seg_words(O, s, 0, SEG_NC<real>, u);
    __builtin_memcpy(&d, &u, seg_core_bytes<real>());
}
template<typename real> __device__ inline void load_seg_rest(const DevOut &O, size_t s, SegDesc<real> &d)
{
    SegWords<real> u;
    constexpr size_t c = seg_core_bytes<real>();
    load_seg_words(O, s, (int)(c / 16), SEG_NW<real>, u);
    __builtin_memcpy(reinterpret_cast<char *>(&d) + c, reinterpret_cast<const char *>(&u) + c,
                     sizeof(SegDesc<real>) - c);
}

template<typename real> struct DevSink {
    const DevOut *O;
    const Prob2D<real> *P;
    unsigned long long nc;
    __device__ void add_field(const Prob2D<real> &Pp, size_t idx, float re, float im)
    {
        unsafeAtomicAdd(Pp.field + 2 * idx, re);
        unsafeAtomicAdd(Pp.field + 2 * idx + 1, im);
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
        c.srcaz = NAN;
        c.rcvaz = NAN;
        if(P->nx_ntheta) { // Nx2D: RcvrAzimAngle = SrcAzimAngle = the bearing (influence.hpp:122-135)
            float az = (float)((RL(180.0) / RC<real>::PI) * P->base[P->o_beta + isz % P->nx_ntheta]);
            c.srcaz = az;
            c.rcvaz = az;
        }
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
        if(P->nx_ntheta) { // Nx2D: itheta = ibeta (influence.hpp:439) and the source indices
            int r = isz / P->nx_ntheta;
            h.itheta = h.ibeta = isz - r * P->nx_ntheta;
            h.isy = r % P->nx_nsy;
            r /= P->nx_nsy;
            h.isx = r % P->nx_nsx;
            h.isz = r / P->nx_nsx;
        }
        O->hits[mi] = h;
    }
    __device__ void count() { ++nc; }
    __device__ void mark() {}
};

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

// Segment slots are reserved SEG_BLOCK at a time per wavefront (one atomic per
// block instead of one per step). Slots of a block that stay unused are marked
// empty (nitems = 0) and skipped by the scatter phase.
constexpr uint32_t SEG_BLOCK = 256;

// The same pointer, marked as global memory: after inlining, the loads and
// stores through it become global_* instructions instead of flat_* ones,
// which the compiler must also count against LDS traffic (every wait for a
// flat load then drains outstanding LDS operations and vice versa).
template<typename T> __device__ inline T *as_global(T *p)
{
    return (T *)((__attribute__((address_space(1))) T *)p);
}

struct SegBlock {
    uint32_t base, used, cap;
};

template<typename real> __device__ inline void seal_block(const SegBlock &B, int lane, const DevOut &O)
{
    for(uint32_t k = B.used + (uint32_t)lane; k < B.cap; k += WAVE) {
        uint32_t slot = B.base + k;
        if(slot < O.seg_cap) clear_seg<real>(O, slot);
    }
}

// Append this lane's descriptor (if it has work items) to the segment buffer:
// all of it, or its core when `full` is false (seg_full). Called by all lanes
// of the wavefront. A slot beyond the buffer sets the overflow flag: the host
// then discards the pass and splits the ray range.
template<typename real>
__device__ inline void append_seg(const SegDesc<real> &d, int lane, const DevOut &O, SegBlock &B, bool full)
{
    bool has = d.nitems > 0;
    unsigned long long m = __ballot(has);
    if(!m) return;
    uint32_t cnt = (uint32_t)__popcll(m);
    if(B.used + cnt > B.cap) {
        seal_block<real>(B, lane, O);
        uint32_t base = 0;
        if(lane == 0) base = atomicAdd(O.seg_count, SEG_BLOCK);
        B.base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        B.used = 0;
        B.cap = SEG_BLOCK;
    }
    if(has) {
        uint32_t slot = B.base + B.used + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if(slot < O.seg_cap) {
            // 16-byte stores through a global-memory pointer (global_store,
            // counted apart from LDS reads; a generic store would make every
            // later LDS wait wait for it too)
            typedef __attribute__((address_space(1))) unsigned long long gu64;
            SegWords<real> u;
            u.d = d;
            gu64 *dst = (gu64 *)O.segs + 2 * (size_t)slot;
            const size_t cap2 = 2 * (size_t)O.seg_cap;
            const int nw = full ? SEG_NW<real> : SEG_NC<real>;
#pragma unroll
            for(int k = 0; k < SEG_NW<real>; ++k)
                if(k < nw) {
                    dst[k * cap2] = u.w[2 * k];
                    dst[k * cap2 + 1] = u.w[2 * k + 1];
                }
        } else {
            atomicOr(O.overflow, 1u);
        }
    }
    B.used += cnt;
}

// Exclusive prefix sum of cnt over the wavefront; returns the total.
__device__ inline int wave_excl_scan(int cnt, int lane, int &excl)
{
    int incl = cnt;
    for(int off = 1; off < WAVE; off <<= 1) {
        int v = __shfl_up(incl, off);
        if(lane >= off) incl += v;
    }
    excl = incl - cnt;
    return __shfl(incl, WAVE - 1);
}

// Lane whose [excl, excl + cnt) holds flattened index t (excl non-decreasing).
__device__ inline int wave_owner(int excl, int t)
{
    int own = 0;
    for(int b = WAVE / 2; b >= 1; b >>= 1) {
        int probe = own + b < WAVE ? own + b : WAVE - 1;
        int e = __shfl(excl, probe);
        if(own + b < WAVE && e <= t) own += b;
    }
    return own;
}

// Scatter one segment's receivers with the 64 lanes of a wavefront (d is
// wavefront-uniform). Both geometries have rows of different lengths --
// Cartesian range columns with their depth bands, ray-centred depths with
// their range windows -- so the items of up to 64 rows are flattened over the
// lanes: a prefix sum of the row lengths, then each lane finds the row of its
// item by a binary search over the lanes' prefix values.
template<typename real, int RT, bool FAST, typename Sink>
__device__ inline void scatter_seg(const SegDesc<real> &d, const Prob2D<real> &P, Sink &sink, int lane)
{
    if(!P.isRayCen) {
        const int nr = (d.ir1 - d.ir0) * d.dir + 1;
        for(int g = 0; g < nr; g += WAVE) {
            int ir = 0, i0 = 0, i1 = -1;
            if(g + lane < nr) cart_column(d, g + lane, P, P.base, ir, i0, i1);
            const int cnt = i1 >= i0 ? i1 - i0 + 1 : 0;
            int excl;
            const int total = wave_excl_scan(cnt, lane, excl);
            for(int t0 = 0; t0 < total; t0 += WAVE) {
                const int t = t0 + lane;
                const int own = wave_owner(excl, t);
                const int oir = __shfl(ir, own), oi0 = __shfl(i0, own), oexcl = __shfl(excl, own);
                if(t < total) item_cart_at<real, RT, Sink, FAST>(d, oir, oi0 + (t - oexcl), P, P.base, P.Rz, sink);
            }
        }
        return;
    }
    const real *Rr = P.base + P.o_Rr;
    for(int g = 0; g < d.nz; g += WAVE) {
        const int iz = d.iz0 + g + lane;
        real nA = 0, rA = 0, nB = 0, rB = 1;
        int blo = 0, bhi = -1;
        if(g + lane < d.nz && !raycen_range(d, iz, P, Rr, P.Rz, nA, rA, nB, rB, blo, bhi)) bhi = blo - 1;
        const int cnt = bhi - blo + 1;
        int excl;
        const int total = wave_excl_scan(cnt, lane, excl);
        for(int t0 = 0; t0 < total; t0 += WAVE) {
            const int t = t0 + lane;
            const int own = wave_owner(excl, t);
            const real onA = __shfl(nA, own), orA = __shfl(rA, own), onB = __shfl(nB, own), orB = __shfl(rB, own);
            const int oblo = __shfl(blo, own), oexcl = __shfl(excl, own), oiz = __shfl(iz, own);
            if(t < total) raycen_one<real, RT, Sink, FAST>(d, oblo + (t - oexcl), oiz, onA, orA, onB, orB, P, Rr, sink);
        }
    }
}

// Trace kernel of the field modes (TL / arrivals / eigen): the ray loop of
// MainFieldModes (src/trace.hpp:525-589). Each step's influence state update
// is done here (prelude); the receiver work goes to the segment buffer.
template<typename real, char ST>
__global__ void __launch_bounds__(256) k_trace2d(Prob2D<real> P, DevOut O)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const real *B = stage(P, lds_raw);
    const real *G = P.base;
    const int lane = threadIdx.x & (WAVE - 1);
    const int njobs = O.job_end - O.job_begin;
    Err e = {0, 0, 0, 0};
    RayPt<real> p0;
    Trace<real> S;
    Infl<real> I;
    int job = 0, isz = 0, ia = 0, is = 0;
    bool active = false, exhausted = false, init_fail = false;
    unsigned long long steps = 0;
    SegBlock blk{0u, 0u, 0u};
    for(int iter = 0;; ++iter) {
        if(!exhausted) {
            unsigned long long need = __ballot(!active);
            if(need) {
                int cnt = __popcll(need);
                int leader = __ffsll((unsigned long long)need) - 1;
                int base = 0;
                if(lane == leader) base = atomicAdd(O.work, cnt);
                base = __shfl(base, leader);
                if(base + cnt >= njobs) exhausted = true;
                if(!active) {
                    int my = base + __popcll(need & ((1ull << lane) - 1ull));
                    if(my < njobs) {
                        job = O.job_begin + my;
                        JobIndices(job, P, isz, ia);
                        is = 0;
                        active = RayInit<real, ST>(isz, ia, p0, S, I, P, B, G, e);
                        init_fail = !active;
                        if(!active && O.stats) {
                            O.stats[3 * job] = 0;
                            O.stats[3 * job + 1] = 0;
                            O.stats[3 * job + 2] = 0;
                            O.xend[2 * job] = 0.0;
                            O.xend[2 * job + 1] = 0.0;
                        }
                    }
                }
                note_done(O.prog, init_fail);
                init_fail = false;
            }
        }
        if(!__any(active)) {
            if(exhausted) break;
            continue;
        }
        if((iter & 31) == 31) {
            // HasErrored (src/util/errors.hpp:48-51), polled every 32 steps and
            // made wavefront-uniform so that all lanes leave together; a full
            // segment buffer ends the pass the same way.
            if(e.err) atomicOr(O.errw + 0, e.err);
            uint32_t ge = __hip_atomic_load(O.errw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t of = __hip_atomic_load(O.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if(__any(ge != 0u || e.err != 0u || of != 0u)) {
                exhausted = true;
                active = false;
                continue;
            }
        }
        RayPt<real> p1, p2;
        SegDesc<real> d;
        d.nitems = 0;
        bool two = false;
        if(active) {
            two = RayUpdate<real, ST>(p0, p1, p2, S, P, B, G, e);
            prelude(d, p0, p1, I, is, P, G, B + P.o_rzs, isz, ia);
            ++is;
        }
        append_seg(d, lane, O, blk, seg_full(P));
        if(__any(two)) {
            d.nitems = 0;
            if(active && two) {
                prelude(d, p1, p2, I, is, P, G, B + P.o_rzs, isz, ia);
                ++is;
            }
            append_seg(d, lane, O, blk, seg_full(P));
        }
        bool fin = false;
        if(active) {
            p0 = two ? p2 : p1;
            int t = RayTerminate(p0, is, S, MAXN, P, e);
            if(t) {
                active = false;
                fin = true;
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
        note_done(O.prog, fin);
    }
    seal_block<real>(blk, lane, O);
    flush_err(e, O);
    for(int off = WAVE / 2; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
    if(lane == 0) atomicAdd(O.counters + 0, steps);
}

// Scatter kernel: one wavefront per segment; the descriptor is wavefront-
// uniform (scalar loads) and the 64 lanes split its receiver work items
// (influence.hpp:1075-1187 ray-centred depths, :1298-1386 Cartesian ranges x
// depths).
template<typename real, int RT, bool FAST>
__global__ void __launch_bounds__(256) k_scatter2d(Prob2D<real> P, DevOut O)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    const int nwaves = gridDim.x * (blockDim.x / WAVE);
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    DevSink<real> sink{&O, &P, 0};
    for(uint32_t s = (uint32_t)wave; s < n; s += (uint32_t)nwaves) {
        SegDesc<real> d;
        load_seg_core(O, s, d);
        if(d.nitems <= 0) continue;
        if(seg_full(P)) load_seg_rest(O, s, d);
        scatter_seg<real, RT, FAST>(d, P, sink, lane);
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) sink.nc += __shfl_xor(sink.nc, off);
    if(lane == 0 && sink.nc) atomicAdd(O.counters + 1, sink.nc);
}

// ---------------------------------------------------------------------------
// Range-binned TL scatter with an LDS field tile.
//
// Segments are counting-sorted by the receiver-range column at their centre
// (bins of BIN_W columns). A workgroup takes a chunk of one bin and keeps the
// complex field of TILE_W columns around the bin (all NRz_per_range depths) in
// LDS; contributions inside the tile are added with LDS atomics, the rest go
// straight to the global field, and the tile is flushed once at the end. With
// ~10^3 contributions per receiver per field, this removes almost all global
// atomics.
// ---------------------------------------------------------------------------
// Single-problem scatter with one lane per segment (segments with few
// receivers each: hat beams on coarse depth windows).
template<typename real, int RT, bool FAST>
__global__ void __launch_bounds__(256) k_scatter2d_lanes(Prob2D<real> P, DevOut O)
{
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    DevSink<real> sink{&O, &P, 0};
    const uint32_t stride = gridDim.x * blockDim.x;
    for(uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
        SegDesc<real> d;
        load_seg_core(O, s, d);
        if(d.nitems <= 0) continue;
        if(seg_full(P)) load_seg_rest(O, s, d);
        seg_items<real, RT, DevSink<real>, FAST>(d, P, P.base, P.Rz, sink);
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) sink.nc += __shfl_xor(sink.nc, off);
    if((threadIdx.x & (WAVE - 1)) == 0 && sink.nc) atomicAdd(O.counters + 1, sink.nc);
}

enum : int { SCATTER_AUTO = 0, SCATTER_TILED = 1, SCATTER_WAVE = 2, SCATTER_LANES = 3 };

struct BinCfg {
    int32_t nbins, bin_w, tile_w, tile_lo_off; // tile columns = [b*bin_w - tile_lo_off, + tile_w)
    int32_t chunk;                              // segments per workgroup
    int32_t *counts, *offsets, *cursor, *chunk_start;
    uint32_t *idx;
    int32_t *nchunks;
};

template<typename real> __device__ inline int seg_bin(const SegDesc<real> &d, const Prob2D<real> &P, const BinCfg &C)
{
    int col;
    if(P.isRayCen) {
        double c = 0.5 * ((double)d.xb.x + (d.lastValid ? (double)d.xa.x : (double)d.xb.x));
        double t = (c - (double)P.base[P.o_Rr]) / (double)P.Delta_r;
        t = t < -1.0 ? -1.0 : (t > (double)P.NRr ? (double)P.NRr : t);
        col = (int)t;
    } else {
        col = (d.ir0 + d.ir1) / 2;
    }
    col = imax(0, imin(col, P.NRr - 1));
    return imin(col / C.bin_w, C.nbins - 1);
}

// The bin of stored segment s (its rest read for ray-centred segments only).
template<typename real> __device__ inline int seg_bin_at(const DevOut &O, size_t s, const Prob2D<real> &P, const BinCfg &C)
{
    SegDesc<real> d;
    load_seg_core(O, s, d);
    if(P.isRayCen) load_seg_rest(O, s, d);
    return seg_bin(d, P, C);
}

// Bin histogram: each workgroup owns a contiguous slice of the segments and
// counts it in LDS first (one global atomic per bin per workgroup).
template<typename real> __global__ void __launch_bounds__(256) k_bin_count(Prob2D<real> P, DevOut O, BinCfg C)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int32_t *h = reinterpret_cast<int32_t *>(lds_raw);
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    uint32_t s0 = blockIdx.x * per, s1 = min(n, s0 + per);
    for(int b = threadIdx.x; b < C.nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for(uint32_t s = s0 + threadIdx.x; s < s1; s += blockDim.x)
        if(seg_nitems<real>(O, s) > 0) atomicAdd(h + seg_bin_at(O, s, P, C), 1);
    __syncthreads();
    for(int b = threadIdx.x; b < C.nbins; b += blockDim.x)
        if(h[b]) atomicAdd(C.counts + b, h[b]);
}

// Exclusive scan of the bin counts and of the per-bin chunk counts (one workgroup).
__global__ void __launch_bounds__(1024) k_bin_scan(BinCfg C)
{
    __shared__ int32_t part_s[1024], part_c[1024];
    int t = threadIdx.x, nb = C.nbins;
    int per = (nb + 1023) / 1024;
    int b0 = t * per, b1 = min(nb, b0 + per);
    int ss = 0, sc = 0;
    for(int b = b0; b < b1; ++b) {
        ss += C.counts[b];
        sc += (C.counts[b] + C.chunk - 1) / C.chunk;
    }
    part_s[t] = ss;
    part_c[t] = sc;
    __syncthreads();
    if(t == 0) {
        int a = 0, c = 0;
        for(int i = 0; i < 1024; ++i) {
            int x = part_s[i], y = part_c[i];
            part_s[i] = a;
            part_c[i] = c;
            a += x;
            c += y;
        }
        *C.nchunks = c;
    }
    __syncthreads();
    ss = part_s[t];
    sc = part_c[t];
    for(int b = b0; b < b1; ++b) {
        C.offsets[b] = ss;
        C.cursor[b] = ss;
        int nc = (C.counts[b] + C.chunk - 1) / C.chunk;
        for(int k = 0; k < nc; ++k) C.chunk_start[sc + k] = b; // chunk -> bin
        ss += C.counts[b];
        sc += nc;
    }
}

// Bin fill with the same slicing as k_bin_count: the workgroup reserves its
// range in every bin with one global atomic, then ranks its segments in LDS.
template<typename real> __global__ void __launch_bounds__(256) k_bin_fill(Prob2D<real> P, DevOut O, BinCfg C)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int32_t *h = reinterpret_cast<int32_t *>(lds_raw);
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    uint32_t s0 = blockIdx.x * per, s1 = min(n, s0 + per);
    for(int b = threadIdx.x; b < C.nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for(uint32_t s = s0 + threadIdx.x; s < s1; s += blockDim.x)
        if(seg_nitems<real>(O, s) > 0) atomicAdd(h + seg_bin_at(O, s, P, C), 1);
    __syncthreads();
    for(int b = threadIdx.x; b < C.nbins; b += blockDim.x)
        if(h[b]) h[b] = atomicAdd(C.cursor + b, h[b]);
    __syncthreads();
    for(uint32_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        if(seg_nitems<real>(O, s) <= 0) continue;
        int pos = atomicAdd(h + seg_bin_at(O, s, P, C), 1);
        C.idx[pos] = s;
    }
}

template<typename real> struct TileSink {
    float *tile;          // LDS [tile_w][nz] complex
    int col0, tile_w, nz, nrr;
    unsigned long long nc;
    __device__ void add_field(const Prob2D<real> &Pp, size_t idx, float re, float im)
    {
        int ir = (int)(idx % (size_t)nrr);
        int iz = (int)((idx / (size_t)nrr) % (size_t)nz);
        int c = ir - col0;
        if(c >= 0 && c < tile_w && (idx / ((size_t)nrr * nz)) == 0) {
            int k = c * nz + iz; // real parts first, then imaginary parts (bank-conflict free)
            atomicAdd(tile + k, re);
            atomicAdd(tile + tile_w * nz + k, im);
        } else {
            unsafeAtomicAdd(Pp.field + 2 * idx, re);
            unsafeAtomicAdd(Pp.field + 2 * idx + 1, im);
        }
    }
    template<typename R2> __device__ void add_arrival(size_t, R2, R2, cpx<R2>, R2, R2, int, int, int, int, int) {}
    __device__ void add_eigen(int, int, int, int, int) {}
    __device__ void count() { ++nc; }
    __device__ void mark() {}
};

// One workgroup per (bin, chunk); the field of source isz lives at
// P.field + isz * plane, so the tile is per source: segments of other
// sources fall back to global atomics (single-source runs are the common case).
constexpr int TILE_THREADS = 512;

template<typename real, bool FAST>
__global__ void __launch_bounds__(TILE_THREADS, 4) k_scatter_tile(Prob2D<real> P, DevOut O, BinCfg C)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float *tile = reinterpret_cast<float *>(lds_raw);
    int nchunks = *C.nchunks;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int nw = blockDim.x / WAVE;
    const int nz = P.NRz_per_range;
    unsigned long long total = 0;
    for(int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        int b = C.chunk_start[ch];
        int first_chunk = ch;
        while(first_chunk > 0 && C.chunk_start[first_chunk - 1] == b) --first_chunk; // chunk index within bin
        int k0 = C.offsets[b] + (ch - first_chunk) * C.chunk;
        int k1 = min(C.offsets[b] + C.counts[b], k0 + C.chunk);
        int col0 = b * C.bin_w - C.tile_lo_off;
        int ncol = C.tile_w;
        for(int i = threadIdx.x; i < 2 * ncol * nz; i += blockDim.x) tile[i] = 0.f;
        __syncthreads();
        TileSink<real> sink{tile, col0, ncol, nz, P.NRr, 0};
        for(int k = k0 + wave; k < k1; k += nw) {
            uint32_t si = __builtin_amdgcn_readfirstlane(C.idx[k]);
            SegDesc<real> d;
            load_seg_core(O, si, d);
            if(seg_full(P)) load_seg_rest(O, si, d);
            if(d.isz != 0) {
                DevSink<real> gs{&O, &P, 0};
                scatter_seg<real, RT_TL, FAST>(d, P, gs, lane);
                sink.nc += gs.nc;
                continue;
            }
            scatter_seg<real, RT_TL, FAST>(d, P, sink, lane);
        }
        __syncthreads();
        // flush: tile column c, depth iz -> field (isz 0)
        for(int i = threadIdx.x; i < ncol * nz; i += blockDim.x) {
            int c = i / nz, iz = i - c * nz;
            int ir = col0 + c;
            if(ir < 0 || ir >= P.NRr) continue;
            float re = tile[i], im = tile[ncol * nz + i];
            if(re != 0.f || im != 0.f) {
                size_t idx = (size_t)iz * P.NRr + ir;
                unsafeAtomicAdd(P.field + 2 * idx, re);
                unsafeAtomicAdd(P.field + 2 * idx + 1, im);
            }
        }
        __syncthreads();
        total += sink.nc;
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) total += __shfl_xor(total, off);
    if(lane == 0 && total) atomicAdd(O.counters + 1, total);
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
    for(int job = blockIdx.x * blockDim.x + threadIdx.x; job < O.job_end; job += gridDim.x * blockDim.x) {
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

// Nx2D trace kernel: the ray loop of MainFieldModes (src/trace.hpp:525-589)
// with O3D && !R3D -- rays on the bearing planes of a 3D environment (nx2d.h),
// influence state and segment descriptors of the 2D beams, so that the 2D
// scatter kernels apply the receivers of every bearing plane. One wavefront
// per workgroup; the 3D environment stays in global memory (L2-resident).
template<typename real>
__global__ void __launch_bounds__(64) k_tracenx2d(Prob2D<real> P, bhc3d::Prob3D<real> N, DevOut O)
{
    using bhc3d::OrgNx;
    using bhc3d::Trace3;
    __shared__ real refl_lds[768];  // the reflection tables in LDS when they fit
    const real *tabs = bhc3d::stage_refl_tables(N, refl_lds, 768);
    const real *G = P.base;
    const int lane = threadIdx.x & (WAVE - 1);
    const int njobs = O.job_end - O.job_begin;
    Err e = {0, 0, 0, 0};
    RayPt<real> p0;
    Trace3<real> S;
    OrgNx<real> org;
    Infl<real> I;
    int job = 0, isf = 0, ia = 0, is = 0;
    bool active = false, exhausted = false, init_fail = false;
    unsigned long long steps = 0;
    SegBlock blk{0u, 0u, 0u};
    for(int iter = 0;; ++iter) {
        if(!exhausted) {
            unsigned long long need = __ballot(!active);
            if(need) {
                int cnt = __popcll(need);
                int leader = __ffsll((unsigned long long)need) - 1;
                int base = 0;
                if(lane == leader) base = atomicAdd(O.work, cnt);
                base = __shfl(base, leader);
                if(base + cnt >= njobs) exhausted = true;
                if(!active) {
                    int my = base + __popcll(need & ((1ull << lane) - 1ull));
                    if(my < njobs) {
                        job = O.job_begin + my;
                        int ib, isx, isy, isz;
                        isf = bhc3d::JobIndicesNx(job, N, ia, ib, isx, isy, isz);
                        is = 0;
                        active = bhc3d::RayInitNx(ia, ib, isx, isy, isz, p0, S, org, I, P, N, e);
                        init_fail = !active;
                        if(!active && O.stats) {
                            O.stats[3 * job] = O.stats[3 * job + 1] = O.stats[3 * job + 2] = 0;
                            O.xend[3 * job] = O.xend[3 * job + 1] = O.xend[3 * job + 2] = 0.0;
                        }
                    }
                }
                note_done(O.prog, init_fail);
                init_fail = false;
            }
        }
        if(!__any(active)) {
            if(exhausted) break;
            continue;
        }
        if((iter & 31) == 31) {
            // HasErrored (src/util/errors.hpp:48-51) and a full segment buffer
            // end the pass for the whole wavefront
            if(e.err) atomicOr(O.errw + 0, e.err);
            uint32_t ge = __hip_atomic_load(O.errw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t of = __hip_atomic_load(O.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if(__any(ge != 0u || e.err != 0u || of != 0u)) {
                exhausted = true;
                active = false;
                continue;
            }
        }
        RayPt<real> p1, p2;
        SegDesc<real> d;
        d.nitems = 0;
        bool two = false;
        if(active) {
            two = bhc3d::RayUpdateNx(p0, p1, p2, S, org, N, P, e, tabs);
            prelude(d, p0, p1, I, is, P, G, G + P.o_rzs, isf, ia);
            ++is;
        }
        append_seg(d, lane, O, blk, seg_full(P));
        if(__any(two)) {
            d.nitems = 0;
            if(active && two) {
                prelude(d, p1, p2, I, is, P, G, G + P.o_rzs, isf, ia);
                ++is;
            }
            append_seg(d, lane, O, blk, seg_full(P));
        }
        bool fin = false;
        if(active) {
            p0 = two ? p2 : p1;
            if(bhc3d::RayTerminateNx(p0, is, S, org, N, e, MAXN)) {
                active = false;
                fin = true;
                steps += (unsigned long long)is;
                if(O.stats) {
                    const bhc3d::v3<real> xe = bhc3d::RayToOceanX(p0.x, org);
                    O.stats[3 * job] = is;
                    O.stats[3 * job + 1] = p0.NumTopBnc;
                    O.stats[3 * job + 2] = p0.NumBotBnc;
                    O.xend[3 * job] = (double)xe.x;
                    O.xend[3 * job + 1] = (double)xe.y;
                    O.xend[3 * job + 2] = (double)xe.z;
                }
            }
        }
        note_done(O.prog, fin);
    }
    seal_block<real>(blk, lane, O);
    flush_err(e, O);
    for(int off = WAVE / 2; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
    if(lane == 0) atomicAdd(O.counters + 0, steps);
}

// Nx2D ray run type: MainRayMode (src/trace.hpp:482-520) with O3D && !R3D.
// Points are in ray coordinates (range along the bearing, depth); org_out
// gets each ray's origin (xs.x, xs.y, cos beta, sin beta).
template<typename real>
__global__ void __launch_bounds__(256) k_raynx2d(Prob2D<real> P, bhc3d::Prob3D<real> N, DevOut O, double *org_out)
{
    Err e = {0, 0, 0, 0};
    const int M = O.max_points;
    for(int job = blockIdx.x * blockDim.x + threadIdx.x; job < O.job_end; job += gridDim.x * blockDim.x) {
        int ia, ib, isx, isy, isz;
        bhc3d::JobIndicesNx(job, N, ia, ib, isx, isy, isz);
        bhc_ray_pt2d *ray = O.points + (size_t)job * M;
        RayPt<real> p0, p1, p2;
        bhc3d::Trace3<real> S;
        bhc3d::OrgNx<real> org;
        Infl<real> I;
        int Nsteps = -1;
        bool ok = bhc3d::RayInitNx(ia, ib, isx, isy, isz, p0, S, org, I, P, N, e);
        store_pt(ray, p0);
        org_out[4 * job] = (double)org.xs.x;
        org_out[4 * job + 1] = (double)org.xs.y;
        org_out[4 * job + 2] = (double)org.tr.x;
        org_out[4 * job + 3] = (double)org.tr.y;
        if(!ok) {
            O.nsteps_out[job] = 1;
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = bhc3d::RayUpdateNx(p0, p1, p2, S, org, N, P, e);
            store_pt(ray + is + 1, p1);
            if(two) store_pt(ray + is + 2, p2);
            is += two ? 2 : 1;
            p0 = two ? p2 : p1;
            int t = bhc3d::RayTerminateNx(p0, is, S, org, N, e, M);
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
            a.SrcAzimAngle = v.srcaz;
            a.RcvrDeclAngle = v.rcv;
            a.RcvrAzimAngle = v.rcvaz;
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
            a.SrcAzimAngle = v.srcaz;
            a.RcvrDeclAngle = v.rcv;
            a.RcvrAzimAngle = v.rcvaz;
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
            a.SrcAzimAngle = w1 * a.SrcAzimAngle + w2 * v.srcaz;
            a.RcvrDeclAngle = w1 * a.RcvrDeclAngle + w2 * v.rcv;
            a.RcvrAzimAngle = w1 * a.RcvrAzimAngle + w2 * v.rcvaz;
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
    hipEvent_t ev0, ev1, evm;   // start, end, trace/scatter boundary of the first pass
    hipEvent_t pa, pb;          // per-pass phase brackets
    double ph_trace_ms, ph_scatter_ms; // trace / scatter kernel time summed over the last run's passes
    int num_cus;
    int lds_max;
    int tpb, bpc;
    char name[256];
    // segment buffer shared by all problems of the context (grown on demand)
    void *segs;
    size_t seg_bytes;
    uint32_t *d_seg_count, *d_overflow;
    uint32_t *d_bin_idx;        // binned segment order (seg capacity entries)
    int passes;                 // trace/scatter passes of the last launch
    int use_tiles;              // 0 disables the LDS-tiled TL scatter
    int exact_phasor;           // 1: TL phasor in the reference's full arithmetic
    size_t seg_mb;              // segment buffer size (MiB), allocated at first use
    double seg_per_ray3;        // 3D runs: largest segments-per-ray seen (sizes the first pass)
    int scatter3_map;           // 3D scatter mapping: 0 one wavefront per segment, 1 one lane per segment,
                                // 2 (default) segment groups flattened per wavefront
    // finished-ray progress of the current or last run (bhc_progress)
    unsigned long long *h_prog; // pinned, device-visible host word written by the trace kernels
    unsigned long long *d_prog; // device counter
    std::atomic<long long> prog_total{0}, prog_seen{0};
    std::atomic<int> prog_final{0};
    // non-blocking runs: one worker thread at a time per context
    std::thread worker;
    std::atomic<int> busy{0};
    int async_rc = 0;
    std::string async_msg;
    // device address of the static scratch array on this context's device
    void *scratch = nullptr;
};

namespace {
// set on a context's worker thread: the run it performs may use the context
thread_local bool tl_async_worker = false;

// Entry points that use a context refuse while a non-blocking run holds it.
int check_idle(bhc_ctx *c, const char *who)
{
    if(!c) return fail("%s: null context", who);
    if(c->busy.load() && !tl_async_worker) return fail("%s: a non-blocking run is in flight on this context", who);
    return 0;
}

// Start of a run of `total` rays: the host word now, the device counter on the
// stream ahead of the run's kernels.
int prog_begin(bhc_ctx *c, long long total)
{
    c->prog_final = 0;
    c->prog_seen = 0;
    c->prog_total = total;
    __atomic_store_n(c->h_prog, 0ull, __ATOMIC_RELAXED);
    HIPCHECK(hipMemsetAsync(c->d_prog, 0, sizeof(unsigned long long), c->stream));
    return 0;
}

Progress prog_of(bhc_ctx *c) { return Progress{c->d_prog, c->h_prog}; }
} // namespace

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
    // range bins of the tiled TL scatter
    int nbins, bin_w, tile_w, tile_lo_off, chunk, max_chunks;
    int scatter_mode;           // SCATTER_AUTO / _TILED / _WAVE / _LANES
    int32_t *d_bins;            // counts | offsets | cursor | chunk_start | nchunks
    // Nx2D runs: the 3D environment the trace kernel reads (P64 / P32 then view
    // the same blob as the receiver side), and 3 reals per ray end point
    int nx2d, xend_dim;
    bhc3d::Prob3D<double> N64;
    bhc3d::Prob3D<float> N32;
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

// Segment buffer of bhc_field_launch_noalloc: a file-scope device array, so
// that launch allocates nothing. Every GPU the library runs on holds its own
// copy until the process ends. In host code the array's name is not a device
// address: host code uses only the address hipGetSymbolAddress returns for
// the current device (kept per device, and in each context).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES];

// Address of `scratch` on `device` (which must be current), looked up once.
int scratch_address(int device, void **ptr)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(device < 0 || device >= MAX_DEVICES) return fail("scratch: device %d out of range", device);
    if(!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p) return fail("scratch: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}
} // namespace

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
    c->use_tiles = 1;
    if(const char *e = getenv("BHC_TL_TILES")) c->use_tiles = atoi(e) != 0;
    c->exact_phasor = 0;
    if(const char *e = getenv("BHC_EXACT_PHASOR")) c->exact_phasor = atoi(e) != 0;
    c->seg_mb = 16384;
    c->scatter3_map = 2;
    if(const char *e = getenv("BHC_SCATTER3_MAP")) c->scatter3_map = std::min(std::max(atoi(e), 0), 2);
    if(const char *e = getenv("BHC_SEG_BUFFER_MB")) c->seg_mb = (size_t)std::max(16L, atol(e));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->ev0));
    HIPCHECK(hipEventCreate(&c->ev1));
    HIPCHECK(hipEventCreate(&c->evm));
    HIPCHECK(hipEventCreate(&c->pa));
    HIPCHECK(hipEventCreate(&c->pb));
    HIPCHECK(hipMalloc((void **)&c->d_seg_count, sizeof(uint32_t)));
    HIPCHECK(hipMalloc((void **)&c->d_overflow, sizeof(uint32_t)));
    HIPCHECK(hipMalloc((void **)&c->d_prog, sizeof(unsigned long long)));
    HIPCHECK(hipHostMalloc((void **)&c->h_prog, 64, hipHostMallocMapped | hipHostMallocCoherent));
    *c->h_prog = 0ull;
    if(scratch_address(device, &c->scratch)) return -1;
    *out = c;
    return 0;
}

extern "C" void bhc_ctx_destroy(bhc_ctx *ctx)
{
    if(!ctx) return;
    if(ctx->worker.joinable()) {
        if(ctx->worker.get_id() == std::this_thread::get_id()) ctx->worker.detach();
        else ctx->worker.join();
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->evm);
    (void)hipEventDestroy(ctx->pa);
    (void)hipEventDestroy(ctx->pb);
    if(ctx->segs) (void)hipFree(ctx->segs);
    if(ctx->d_bin_idx) (void)hipFree(ctx->d_bin_idx);
    (void)hipFree(ctx->d_seg_count);
    (void)hipFree(ctx->d_overflow);
    (void)hipFree(ctx->d_prog);
    (void)hipHostFree(ctx->h_prog);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int bhc_progress(bhc_ctx *ctx, int64_t *done, int64_t *total)
{
    if(!ctx || !done || !total) return fail("bhc_progress: null argument");
    const long long t = ctx->prog_total.load();
    long long d = t;
    if(!ctx->prog_final.load()) {
        d = std::min((long long)__atomic_load_n(ctx->h_prog, __ATOMIC_RELAXED), t);
        long long seen = ctx->prog_seen.load();
        while(d > seen && !ctx->prog_seen.compare_exchange_weak(seen, d)) {
        }
        d = std::max(d, seen);
    }
    *done = d;
    *total = t;
    return 0;
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

namespace {
// Plain streaming copy (16 B per lane per iteration): the known-size kernel
// against which profiler byte counters (FETCH_SIZE, WRITE_SIZE) are scaled.
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for(size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
} // namespace

extern "C" int bhc_copy_calibrate(bhc_ctx *ctx, long long bytes, double *ms)
{
    if(check_idle(ctx, "bhc_copy_calibrate")) return -1;
    if(bytes < 16 || (bytes & 15)) return fail("bhc_copy_calibrate: bytes must be a positive multiple of 16");
    HIPCHECK(hipSetDevice(ctx->device));
    void *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc(&a, (size_t)bytes));
    if(hipMalloc(&b, (size_t)bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail("bhc_copy_calibrate: allocation failed");
    }
    int rc = 0;
    const size_t n = (size_t)bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 16);
    if(hipMemsetAsync(a, 1, (size_t)bytes, ctx->stream) != hipSuccess) rc = fail("bhc_copy_calibrate: memset failed");
    if(!rc) {
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, ctx->stream, (const uint4 *)a, (uint4 *)b, n);
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        float f = 0.f;
        if(hipGetLastError() != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess
           || hipEventElapsedTime(&f, ctx->ev0, ctx->ev1) != hipSuccess)
            rc = fail("bhc_copy_calibrate: copy kernel failed");
        if(ms) *ms = f;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

extern "C" int bhc_last_phase_ms(bhc_ctx *ctx, double *trace_ms, double *scatter_ms, int *passes)
{
    HIPCHECK(hipEventSynchronize(ctx->ev1));
    *trace_ms = ctx->ph_trace_ms;
    *scatter_ms = ctx->ph_scatter_ms;
    *passes = ctx->passes;
    return 0;
}

// Per-pass phase timing: a trace (or scatter) launch is bracketed by pa / pb on
// the context's stream and its duration added once the pass has finished.
static int phase_add(bhc_ctx *c, double *acc)
{
    float ms = 0.f;
    HIPCHECK(hipEventSynchronize(c->pb));
    HIPCHECK(hipEventElapsedTime(&ms, c->pa, c->pb));
    *acc += ms;
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
    void *ptrs[] = {dp->d_bins, dp->d_blob, dp->d_Rz,   dp->d_Sz,     dp->d_field,    dp->d_cand, dp->d_ncand,
                    dp->d_arr,  dp->d_narr, dp->d_hits,   dp->d_neigen,   dp->d_stats, dp->d_xend,
                    dp->d_counters, dp->d_errw, dp->d_work};
    for(void *p : ptrs)
        if(p) (void)hipFree(p);
    delete dp;
}

namespace {
int ensure_segs(bhc_ctx *c);
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
    dp->xend_dim = 2;
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
    // Stage in LDS when the SSP + boundary tables fit the workgroup's LDS.
    dp->lds_bytes = (staged <= (size_t)std::min(ctx->lds_max, 65536)) ? staged : 0;
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
    dp->P64.field = dp->P32.field = dp->d_field;
    dp->P64.njobs = dp->P32.njobs = dp->njobs;
    if(tl && prob->NRr > 1 && prob->Delta_r > 0) {
        // LDS tile of 64 KiB: tile_w range columns x all depths, complex float
        int nz = prob->NRz_per_range;
        int tile_w = (int)(65536 / ((size_t)nz * 2 * sizeof(float)));
        int margin = (int)std::ceil(prob->deltas / (2.0 * prob->Delta_r)) + 2;
        if(tile_w >= 6 && (prob->NRr + std::max(1, tile_w - 2 * margin) - 1) / std::max(1, tile_w - 2 * margin) <= 16384) {
            int bin_w = std::max(1, tile_w - 2 * margin);
            if(bin_w == 1) margin = (tile_w - 1) / 2;
            dp->tile_w = tile_w;
            dp->bin_w = bin_w;
            dp->tile_lo_off = margin;
            dp->nbins = (prob->NRr + bin_w - 1) / bin_w;
            dp->chunk = 2048;
            if(ensure_segs(ctx)) {
                bhc_problem_release(dp);
                return -1;
            }
            size_t cap = ctx->seg_bytes / sizeof(Seg<float>);
            dp->max_chunks = (int)std::min<size_t>(cap / dp->chunk + dp->nbins + 1, 0x7FFFFFFF);
            if(dalloc(&dp->d_bins, 3 * (size_t)dp->nbins + 1 + dp->max_chunks)) {
                bhc_problem_release(dp);
                return -1;
            }
        }
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
    O.job_begin = 0;
    O.job_end = dp->njobs;
    O.prog = prog_of(dp->ctx);
    return O;
}

size_t seg_size(int precision) { return precision == 64 ? sizeof(Seg<double>) : sizeof(Seg<float>); }

// Segment buffer: BHC_SEG_BUFFER_MB (default 16384 MiB) of HBM, allocated once per context.
int ensure_segs(bhc_ctx *c)
{
    if(c->segs) return 0;
    c->seg_bytes = c->seg_mb << 20;
    HIPCHECK(hipMalloc(&c->segs, c->seg_bytes));
    size_t maxseg = c->seg_bytes / sizeof(Seg<float>);
    HIPCHECK(hipMalloc((void **)&c->d_bin_idx, maxseg * sizeof(uint32_t)));
    return 0;
}

template<typename real> const bhc3d::Prob3D<real> &nx_of(const bhc_dev_problem *dp);
template<> const bhc3d::Prob3D<double> &nx_of<double>(const bhc_dev_problem *dp) { return dp->N64; }
template<> const bhc3d::Prob3D<float> &nx_of<float>(const bhc_dev_problem *dp) { return dp->N32; }

// Trace / scatter passes over the problem's rays. segbuf == null: the
// context's segment buffer (allocated at first use); otherwise the given
// buffer of segbytes, with the scatter restricted to the mappings that need
// no range bins (the bins are allocated with the context's buffer).
template<typename real, int RT>
int run_passes(bhc_dev_problem *dp, const Prob2D<real> &P, DevOut O, void *segbuf = nullptr, size_t segbytes = 0)
{
    bhc_ctx *c = dp->ctx;
    hipStream_t s = c->stream;
    const bool own_buffer = segbuf != nullptr;
    if(!own_buffer) {
        if(ensure_segs(c)) return -1;
        segbuf = c->segs;
        segbytes = c->seg_bytes;
    }
    O.segs = segbuf;
    O.seg_count = c->d_seg_count;
    O.overflow = c->d_overflow;
    O.seg_cap = (uint32_t)std::min<size_t>(segbytes / sizeof(Seg<real>), 0xFFFFFFF0u);
    std::vector<std::pair<int, int>> todo{{0, dp->njobs}};
    c->passes = 0;
    c->ph_trace_ms = c->ph_scatter_ms = 0.0;
    bool first = true;
    unsigned long long steps_done = 0;
    while(!todo.empty()) {
        auto r = todo.back();
        todo.pop_back();
        O.job_begin = r.first;
        O.job_end = r.second;
        HIPCHECK(hipMemsetAsync(c->d_seg_count, 0, sizeof(uint32_t), s));
        HIPCHECK(hipMemsetAsync(c->d_overflow, 0, sizeof(uint32_t), s));
        HIPCHECK(hipMemsetAsync(dp->d_work, 0, sizeof(int32_t), s));
        int nj = r.second - r.first;
        int waves = (nj + WAVE - 1) / WAVE;
        // one wavefront per workgroup spreads the rays over all CUs (the trace
        // kernel is register-bound at one wavefront per SIMD)
        int tblocks = std::max(1, std::min(waves, c->num_cus * 4 * c->bpc));
        // error / warning words as the completed passes left them (stream
        // order: after the previous pass's scatter), restored if this trace
        // pass overflows and its rays are traced again
        uint32_t ew_start[4] = {0u, 0u, 0u, 0u};
        unsigned long long prog_start = 0;
        HIPCHECK(hipMemcpyAsync(ew_start, dp->d_errw, sizeof(ew_start), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(&prog_start, c->d_prog, sizeof(prog_start), hipMemcpyDeviceToHost, s));
        if(first) HIPCHECK(hipEventRecord(c->ev0, s));
        HIPCHECK(hipEventRecord(c->pa, s));
        if(dp->nx2d) {
            hipLaunchKernelGGL((k_tracenx2d<real>), dim3(tblocks), dim3(WAVE), 0, s, P, nx_of<real>(dp), O);
        } else switch(P.ssp_type) {
        case 'C': hipLaunchKernelGGL((k_trace2d<real, 'C'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        case 'S': hipLaunchKernelGGL((k_trace2d<real, 'S'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        case 'P': hipLaunchKernelGGL((k_trace2d<real, 'P'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        case 'N': hipLaunchKernelGGL((k_trace2d<real, 'N'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        case 'Q': hipLaunchKernelGGL((k_trace2d<real, 'Q'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        default: hipLaunchKernelGGL((k_trace2d<real, 'A'>), dim3(tblocks), dim3(WAVE), dp->lds_bytes, s, P, O); break;
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(c->pb, s));
        uint32_t of = 0;
        unsigned long long steps_now = 0;
        HIPCHECK(hipMemcpyAsync(&of, c->d_overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(&steps_now, dp->d_counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if(phase_add(c, &c->ph_trace_ms)) return -1;
        if(of) {
            if(nj <= 1) return fail("segment buffer too small for a single ray (raise BHC_SEG_BUFFER_MB)");
            int mid = r.first + nj / 2;
            todo.push_back({mid, r.second});
            todo.push_back({r.first, mid});
            // the trace pass wrote only per-ray statistics, the step counter and
            // the error / warning words: the retried halves rewrite the
            // statistics; the counter and the words go back to the completed
            // passes' values, so that the retried rays' warnings count once
            HIPCHECK(hipMemcpyAsync(dp->d_counters, &steps_done, sizeof(unsigned long long), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(dp->d_errw, ew_start, sizeof(ew_start), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(c->d_prog, &prog_start, sizeof(prog_start), hipMemcpyHostToDevice, s));
            HIPCHECK(hipStreamSynchronize(s));
            continue;
        }
        steps_done = steps_now;
        if(first) HIPCHECK(hipEventRecord(c->evm, s));
        HIPCHECK(hipEventRecord(c->pa, s));
        int mode = dp->scatter_mode;
        if(mode == SCATTER_AUTO) mode = (RT == RT_TL && c->use_tiles && dp->nbins > 0) ? SCATTER_TILED : SCATTER_WAVE;
        if(mode == SCATTER_TILED && (!(RT == RT_TL && dp->nbins > 0) || own_buffer)) mode = SCATTER_WAVE;
        if(mode == SCATTER_LANES) {
            int sblocks = c->num_cus * 8;
            if(RT == RT_TL && !c->exact_phasor)
                hipLaunchKernelGGL((k_scatter2d_lanes<real, RT, true>), dim3(sblocks), dim3(256), 0, s, P, O);
            else
                hipLaunchKernelGGL((k_scatter2d_lanes<real, RT, false>), dim3(sblocks), dim3(256), 0, s, P, O);
        } else if(mode == SCATTER_TILED) {
            BinCfg B;
            B.nbins = dp->nbins;
            B.bin_w = dp->bin_w;
            B.tile_w = dp->tile_w;
            B.tile_lo_off = dp->tile_lo_off;
            B.chunk = dp->chunk;
            B.counts = dp->d_bins;
            B.offsets = B.counts + dp->nbins;
            B.cursor = B.offsets + dp->nbins;
            B.nchunks = B.cursor + dp->nbins;
            B.chunk_start = B.nchunks + 1;
            B.idx = c->d_bin_idx;
            HIPCHECK(hipMemsetAsync(B.counts, 0, dp->nbins * sizeof(int32_t), s));
            int gb = c->num_cus * 4;
            size_t hl = (size_t)dp->nbins * sizeof(int32_t);
            hipLaunchKernelGGL((k_bin_count<real>), dim3(gb), dim3(256), hl, s, P, O, B);
            hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, B);
            hipLaunchKernelGGL((k_bin_fill<real>), dim3(gb), dim3(256), hl, s, P, O, B);
            size_t lds = (size_t)dp->tile_w * dp->NRz_per_range * 2 * sizeof(float);
            if(c->exact_phasor)
                hipLaunchKernelGGL((k_scatter_tile<real, false>), dim3(c->num_cus * 4), dim3(TILE_THREADS), lds, s,
                                   P, O, B);
            else
                hipLaunchKernelGGL((k_scatter_tile<real, true>), dim3(c->num_cus * 4), dim3(TILE_THREADS), lds, s,
                                   P, O, B);
        } else {
            int sblocks = c->num_cus * 8;
            if(RT == RT_TL && !c->exact_phasor)
                hipLaunchKernelGGL((k_scatter2d<real, RT, true>), dim3(sblocks), dim3(256), 0, s, P, O);
            else
                hipLaunchKernelGGL((k_scatter2d<real, RT, false>), dim3(sblocks), dim3(256), 0, s, P, O);
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(c->pb, s));
        if(phase_add(c, &c->ph_scatter_ms)) return -1;
        ++c->passes;
        first = false;
    }
    HIPCHECK(hipEventRecord(c->ev1, s));
    return 0;
}
} // namespace

extern "C" int bhc_problem_autotune(bhc_ctx *ctx, bhc_dev_problem *dp, int *mode_out);

namespace {
int field_launch(bhc_ctx *ctx, bhc_dev_problem *dp, void *segbuf, size_t segbytes)
{
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if(prog_begin(ctx, dp->njobs)) return -1;
    if(dp->d_field) HIPCHECK(hipMemsetAsync(dp->d_field, 0, 2 * dp->nrcv * sizeof(float), s));
    if(dp->d_ncand) HIPCHECK(hipMemsetAsync(dp->d_ncand, 0, dp->nrcv * sizeof(int32_t), s));
    if(dp->d_narr) HIPCHECK(hipMemsetAsync(dp->d_narr, 0, dp->nrcv * sizeof(int32_t), s));
    if(dp->d_neigen) HIPCHECK(hipMemsetAsync(dp->d_neigen, 0, sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(dp->d_counters, 0, 2 * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(dp->d_errw, 0, 4 * sizeof(uint32_t), s));
    DevOut O = make_out(dp);
    char r0 = dp->run0;
    int rt = (r0 == 'C' || r0 == 'S' || r0 == 'I') ? RT_TL : (r0 == 'E') ? RT_EIG : RT_ARR;
    int rc;
    if(dp->precision == 64) {
        rc = rt == RT_TL ? run_passes<double, RT_TL>(dp, dp->P64, O, segbuf, segbytes)
            : rt == RT_ARR ? run_passes<double, RT_ARR>(dp, dp->P64, O, segbuf, segbytes)
                           : run_passes<double, RT_EIG>(dp, dp->P64, O, segbuf, segbytes);
    } else {
        rc = rt == RT_TL ? run_passes<float, RT_TL>(dp, dp->P32, O, segbuf, segbytes)
            : rt == RT_ARR ? run_passes<float, RT_ARR>(dp, dp->P32, O, segbuf, segbytes)
                           : run_passes<float, RT_EIG>(dp, dp->P32, O, segbuf, segbytes);
    }
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
    ctx->prog_final = 1;   // every trace pass has completed: all rays have finished
    return 0;
}
} // namespace

extern "C" int bhc_field_launch(bhc_ctx *ctx, bhc_dev_problem *dp)
{
    if(check_idle(ctx, "bhc_field_launch")) return -1;
    if(!dp || dp->ctx != ctx) return fail("bhc_field_launch: problem not uploaded on this context");
    return field_launch(ctx, dp, nullptr, 0);
}

extern "C" int bhc_field_launch_noalloc(bhc_ctx *ctx, bhc_dev_problem *dp)
{
    if(check_idle(ctx, "bhc_field_launch_noalloc")) return -1;
    if(!dp || dp->ctx != ctx) return fail("bhc_field_launch_noalloc: problem not uploaded on this context");
    if(dp->run0 != 'C' && dp->run0 != 'S' && dp->run0 != 'I')
        return fail("bhc_field_launch_noalloc: TL run types (C, S, I) only, not '%c'", dp->run0);
    if(!ctx->scratch) return fail("bhc_field_launch_noalloc: the scratch array is not mapped on device %d", ctx->device);
    // a pass must hold one ray's segments (at most MAXN, plus a partly used
    // slot block at either end); more rays are split over further passes
    const size_t need = ((size_t)MAXN + 2 * (size_t)SEG_BLOCK) * seg_size(dp->precision);
    if(need > NOALLOC_SCRATCH_BYTES)
        return fail("bhc_field_launch_noalloc: one pass needs %zu bytes of scratch, the static scratch holds %zu",
                    need, NOALLOC_SCRATCH_BYTES);
    return field_launch(ctx, dp, ctx->scratch, NOALLOC_SCRATCH_BYTES);
}

extern "C" int bhc_problem_set_scatter(bhc_dev_problem *dp, int mode)
{
    if(!dp || mode < SCATTER_AUTO || mode > SCATTER_LANES) return fail("bhc_problem_set_scatter: bad arguments");
    dp->scatter_mode = mode;
    return 0;
}

// Scatter-mapping search for a resident problem: every available mapping
// (range-binned LDS tiles for TL runs, one wavefront per segment, one lane per
// segment) is launched twice and the fastest kept for later launches.
extern "C" int bhc_problem_autotune(bhc_ctx *ctx, bhc_dev_problem *dp, int *mode_out)
{
    if(check_idle(ctx, "bhc_problem_autotune")) return -1;
    if(!dp || dp->ctx != ctx) return fail("bhc_problem_autotune: problem not uploaded on this context");
    int modes[3] = {SCATTER_TILED, SCATTER_WAVE, SCATTER_LANES};
    int best_mode = SCATTER_AUTO;
    double best = 1e300;
    bool tl = dp->run0 == 'C' || dp->run0 == 'S' || dp->run0 == 'I';
    for(int m : modes) {
        if(m == SCATTER_TILED && !(tl && dp->nbins > 0 && ctx->use_tiles)) continue;
        dp->scatter_mode = m;
        double t = 0.0;
        for(int rep = 0; rep < 2; ++rep) {
            if(bhc_field_launch(ctx, dp)) return -1;
            double ms = 0.0;
            if(bhc_last_kernel_ms(ctx, &ms)) return -1;
            if(rep == 1) t = ms;
        }
        if(t < best) {
            best = t;
            best_mode = m;
        }
    }
    dp->scatter_mode = best_mode;
    if(mode_out) *mode_out = best_mode;
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
            HIPCHECK(hipMemcpy(out->ray_xend, dp->d_xend, (size_t)dp->xend_dim * dp->njobs * sizeof(double),
                               hipMemcpyDeviceToHost));
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
    if(check_idle(ctx, "bhc_field_run2d")) return -1;
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
    if(check_idle(ctx, "bhc_ray_run2d")) return -1;
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
    if(prog_begin(ctx, njobs)) {
        bhc_problem_release(dp);
        return -1;
    }
    DevOut O = make_out(dp);
    O.job_begin = 0;
    O.job_end = njobs;
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
    if(!rc) ctx->prog_final = 1;
    return rc;
}

// ===========================================================================
// Nx2D runs (bellhopcxx Nx2D, RunType[5] == '2'): 2D rays on the bearing
// planes of a 3D environment. The receiver side -- field, arrivals, eigen
// hits, range bins, per-ray statistics, the trace / scatter passes and the
// arrivals merge -- is the 2D run's over NSz * NSx * NSy * Ntheta source
// slots (one per bearing plane); only the trace kernel differs.
// ===========================================================================
namespace {
int upload_nx2d(bhc_ctx *ctx, const bhc_problem3d *p, const bhc_field_out *spec, char run0, bhc_dev_problem **out)
{
    *out = nullptr;
    if(!p) return fail("null problem");
    if(p->RunType[5] != '2') return fail("not an Nx2D problem (RunType[5] must be '2')");
    if(p->precision != 64 && p->precision != 32) return fail("precision must be 32 or 64");
    if(p->NSx < 1 || p->NSy < 1 || p->NSz < 1 || p->NRz < 1 || p->NRr < 1 || p->Ntheta < 1 || p->Nalpha < 1
       || p->Nbeta < 1)
        return fail("Nx2D problem: sources, receivers and launch angles must be non-empty");
    if(p->RunType[4] == 'I') return fail("Nx2D problem: irregular receiver grids are not supported");
    if(run0 != 'R' && p->Nbeta != p->Ntheta)
        return fail("Nx2D problem: the beams must lie on the receiver bearings (Nbeta == Ntheta)");
    if(p->iSingle_beta < 0 || p->iSingle_beta > p->Nbeta) return fail("iSingle_beta out of range");
    const long long nsrc = (long long)p->NSz * p->NSx * p->NSy * p->Ntheta;
    const long long nj = (long long)(p->iSingle_alpha >= 1 ? 1 : p->Nalpha) * (p->iSingle_beta >= 1 ? 1 : p->Nbeta)
        * p->NSx * p->NSy * p->NSz;
    if(nsrc * p->Nalpha >= (1ll << 30) || nj >= (1ll << 30)) return fail("Nx2D problem: too many rays");
    // the 2D stand-in that sizes the receiver side: nsrc source slots, the
    // problem's receivers, declinations and beam pattern; its profile and
    // boundaries are placeholders (the trace kernel reads the 3D ones)
    static const double z2[2] = {0.0, 1.0}, c2[4] = {1500.0, 0.0, 1500.0, 0.0}, zero4[4] = {0, 0, 0, 0};
    static const double rho2[2] = {1.0, 1.0};
    bhc_bdry_node2d nodes[2];
    memset(nodes, 0, sizeof(nodes));
    nodes[0].x[0] = -1.0e10;
    nodes[1].x[0] = 1.0e10;
    nodes[0].t[0] = nodes[1].t[0] = 1.0;
    nodes[0].n[1] = nodes[1].n[1] = 1.0;
    std::vector<float> sz((size_t)nsrc, 0.0f);
    for(int i = 0; i < p->NSz; ++i) sz[i] = p->Sz[i];
    bhc_problem2d q;
    memset(&q, 0, sizeof(q));
    q.precision = p->precision;
    memcpy(q.RunType, p->RunType, 7);
    q.RunType[0] = run0;
    q.RunType[5] = ' ';
    memcpy(q.BeamType, p->BeamType, 4);
    q.BeamType[0] = p->RunType[1];
    q.freq0 = p->freq0;
    q.ssp_type = 'C';
    q.ssp_npts = 2;
    q.ssp_z = z2;
    q.ssp_c = c2;
    q.ssp_cz = zero4;
    q.ssp_rho = rho2;
    q.top_hs = p->top_hs;
    q.bot_hs = p->bot_hs;
    q.ati_type[0] = q.bty_type[0] = 'L';
    q.ati_type[1] = q.bty_type[1] = 'S';
    q.ati_npts = q.bty_npts = 2;
    q.ati = nodes;
    q.bty = nodes;
    q.trc_npts = p->trc_npts;
    q.brc_npts = p->brc_npts;
    q.trc = p->trc;
    q.brc = p->brc;
    q.NSz = (int32_t)nsrc;
    q.NRz = p->NRz;
    q.NRr = p->NRr;
    q.NRz_per_range = p->NRz;
    q.Sz = sz.data();
    q.Rz = p->Rz;
    q.Rr = p->Rr;
    q.Delta_r = p->Delta_r;
    q.Nalpha = p->Nalpha;
    q.iSingle_alpha = p->iSingle_alpha;
    q.alpha = p->alpha;
    q.alpha_cos = p->alpha_cos;
    q.alpha_sin = p->alpha_sin;
    q.Dalpha = p->Dalpha;
    q.deltas = p->deltas;
    q.box_r = p->box_x;
    q.box_z = p->box_z;
    q.NSBPPts = p->NSBPPts;
    q.sbp = p->sbp;
    bhc_dev_problem *dp = nullptr;
    if(bhc_problem_upload2d(ctx, &q, spec, &dp)) return -1;
    dp->njobs = (int)nj;
    dp->nx2d = 1;
    dp->xend_dim = 3;
    int rc = 0;
    if(dp->d_xend) {
        (void)hipFree(dp->d_xend);
        dp->d_xend = nullptr;
        rc = dalloc(&dp->d_xend, 3 * (size_t)nj);
    }
    // the Nx2D blob replaces the stand-in's
    (void)hipFree(dp->d_blob);
    dp->d_blob = nullptr;
    std::string err;
    if(!rc && p->precision == 64) {
        std::vector<double> blob;
        rc = bhc3d::pack_nx<double>(p, blob, dp->N64, dp->P64, err);
        if(rc) fail("%s", err.c_str());
        if(!rc) rc = dalloc((double **)&dp->d_blob, blob.size());
        if(!rc && hipMemcpy(dp->d_blob, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail("blob upload failed");
    } else if(!rc) {
        std::vector<float> blob;
        rc = bhc3d::pack_nx<float>(p, blob, dp->N32, dp->P32, err);
        if(rc) fail("%s", err.c_str());
        if(!rc) rc = dalloc((float **)&dp->d_blob, blob.size());
        if(!rc && hipMemcpy(dp->d_blob, blob.data(), blob.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail("blob upload failed");
    }
    if(rc) {
        bhc_problem_release(dp);
        return -1;
    }
    dp->lds_bytes = 0;
    dp->P64.base = dp->N64.base = dp->N64.s.base = (const double *)dp->d_blob;
    dp->P32.base = dp->N32.base = dp->N32.s.base = (const float *)dp->d_blob;
    dp->P64.Rz = dp->P32.Rz = dp->N64.Rz = dp->N32.Rz = dp->d_Rz;
    dp->P64.Sz = dp->P32.Sz = dp->N64.Sz = dp->N32.Sz = dp->d_Sz;
    dp->P64.field = dp->P32.field = dp->N64.field = dp->N32.field = dp->d_field;
    dp->P64.njobs = dp->P32.njobs = dp->N64.njobs = dp->N32.njobs = (int)nj;
    *out = dp;
    return 0;
}
} // namespace

extern "C" int bhc_field_runnx2d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out)
{
    if(check_idle(ctx, "bhc_field_runnx2d")) return -1;
    HIPCHECK(hipSetDevice(ctx->device));
    const char r0 = prob ? prob->RunType[0] : 0;
    if(r0 != 'C' && r0 != 'S' && r0 != 'I' && r0 != 'A' && r0 != 'a' && r0 != 'E' && r0 != 'V' && r0 != 'v')
        return fail("bhc_field_runnx2d: run type '%c' is not a field run", r0);
    if(!out) return fail("bhc_field_runnx2d: null output");
    const char r1 = prob->RunType[1];
    if((r1 == 'g' || r1 == 'b') && prob->NRr > 1 && !(prob->Delta_r > 0))
        return fail("bhc_field_runnx2d: ray-centred beams need equally spaced receiver ranges");
    bhc_dev_problem *dp = nullptr;
    if(upload_nx2d(ctx, prob, out, r0, &dp)) return -1;
    int rc = field_launch(ctx, dp, nullptr, 0);
    if(!rc) rc = bhc_field_download(ctx, dp, out);
    bhc_problem_release(dp);
    return rc;
}

extern "C" int bhc_ray_runnx2d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_ray_out *out, double *origin)
{
    if(check_idle(ctx, "bhc_ray_runnx2d")) return -1;
    if(!out || !out->points || !out->nsteps || !origin) return fail("bhc_ray_runnx2d: null output");
    if(out->max_points_per_ray < 3) return fail("max_points_per_ray must be >= 3");
    HIPCHECK(hipSetDevice(ctx->device));
    bhc_field_out spec;
    memset(&spec, 0, sizeof(spec));
    bhc_dev_problem *dp = nullptr;
    if(upload_nx2d(ctx, prob, &spec, 'R', &dp)) return -1;
    const int njobs = dp->njobs;
    if(prog_begin(ctx, njobs)) {
        bhc_problem_release(dp);
        return -1;
    }
    DevOut O = make_out(dp);
    O.job_begin = 0;
    O.job_end = njobs;
    O.max_points = out->max_points_per_ray;
    int32_t *d_nout = nullptr;
    double *d_org = nullptr;
    bhc_ray_pt2d *d_pts = nullptr;
    int rc = 0;
    const size_t npts = (size_t)njobs * out->max_points_per_ray;
    if(dalloc(&d_pts, npts) || dalloc(&d_nout, njobs) || dalloc(&d_org, 4 * (size_t)njobs)) rc = -1;
    if(!rc) {
        O.points = d_pts;
        O.nsteps_out = d_nout;
        hipMemsetAsync(dp->d_errw, 0, 4 * sizeof(uint32_t), ctx->stream);
        hipMemsetAsync(d_pts, 0, npts * sizeof(bhc_ray_pt2d), ctx->stream);
        int blocks = std::max(1, std::min((njobs + 255) / 256, ctx->num_cus * ctx->bpc));
        if(dp->precision == 64)
            hipLaunchKernelGGL(k_raynx2d<double>, dim3(blocks), dim3(256), 0, ctx->stream, dp->P64, dp->N64, O, d_org);
        else
            hipLaunchKernelGGL(k_raynx2d<float>, dim3(blocks), dim3(256), 0, ctx->stream, dp->P32, dp->N32, O, d_org);
        if(hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail("Nx2D ray kernel failed");
        } else {
            hipMemcpy(out->points, d_pts, npts * sizeof(bhc_ray_pt2d), hipMemcpyDeviceToHost);
            hipMemcpy(out->nsteps, d_nout, njobs * sizeof(int32_t), hipMemcpyDeviceToHost);
            hipMemcpy(origin, d_org, 4 * (size_t)njobs * sizeof(double), hipMemcpyDeviceToHost);
            uint32_t ew[4];
            hipMemcpy(ew, dp->d_errw, sizeof(ew), hipMemcpyDeviceToHost);
            out->err = ew[0];
            out->warn = ew[1];
            out->err_count = ew[2];
            out->warn_count = ew[3];
        }
    }
    void *fr[] = {d_nout, d_pts, d_org};
    for(void *q : fr)
        if(q) (void)hipFree(q);
    bhc_problem_release(dp);
    if(!rc) ctx->prog_final = 1;
    return rc;
}

// ===========================================================================
// Batched sweeps: many independent 2D TL problems in one persistent launch.
//
// Every wavefront works on one problem at a time (its descriptor is loaded
// with scalar loads and its SSP / boundary tables are staged in the
// workgroup's LDS), its lanes take that problem's rays from a per-problem
// work counter as they free up, and when the problem has no unclaimed rays
// left and the wavefront's rays have all ended, the wavefront moves to the
// first problem that still has rays. Segments carry their problem index and
// one scatter pass serves every problem. When the segment buffer nears its
// end, every wavefront saves its in-flight rays and exits; the host scatters
// what was traced and relaunches, and the saved rays resume where they
// stopped. The SSP type is dispatched at run time.
// ===========================================================================
namespace {

template<typename real> struct RaySave {
    RayPt<real> p0;
    Trace<real> S;
    Infl<real> I;
    int32_t job, is, isz, ia, active, pad;
};

template<typename real> struct BatchDev {
    const Prob2D<real> *probs;
    int32_t nprob;
    int32_t lds_cap;               // reals of LDS per workgroup for staging
    int32_t *ray_next;             // [nprob] next unclaimed ray
    int32_t *pcursor;              // lowest problem that may have unclaimed rays
    uint32_t *perr;                // [nprob][4] err, warn, nerr, nwarn
    unsigned long long *psteps;    // [nprob] steps
    const RaySave<real> *save_in;  // [*nsave_in][64]
    const int32_t *save_in_pid;
    const int32_t *nsave_in;       // saved wavefronts of the previous pass
    int32_t *resume_next;
    RaySave<real> *save_out;
    int32_t *save_out_pid;
    int32_t *nsave_out;
    int32_t *more;                 // set when a wavefront stopped taking new problems
    uint32_t suspend_at;           // segment count at which wavefronts suspend
};

template<typename real> __device__ inline const real *stage_problem(const Prob2D<real> &P, int lds_cap,
                                                                    unsigned char *lds)
{
    if(P.staged_reals <= 0 || P.staged_reals > lds_cap) return P.base;
    real *L = reinterpret_cast<real *>(lds);
    __syncthreads();
    for(int i = threadIdx.x; i < P.staged_reals; i += blockDim.x) L[i] = P.base[i];
    __syncthreads();
    return L;
}

__device__ inline void flush_problem_err(Err &e, uint32_t *perr)
{
    if(e.err) atomicOr(perr + 0, e.err);
    if(e.warn) atomicOr(perr + 1, e.warn);
    if(e.nerr) atomicAdd(perr + 2, e.nerr);
    if(e.nwarn) atomicAdd(perr + 3, e.nwarn);
    e = Err{0, 0, 0, 0};
}

// How a batched trace hands each step to the receivers: as a descriptor in
// the segment buffer for a later scatter pass, or fused (each lane evaluates
// its own step's receivers at once, with the fast or the reference-arithmetic
// phasor), which needs no segment buffer and no second kernel.
enum { TRACE_DESC = 0, TRACE_FUSED_FAST = 1, TRACE_FUSED_EXACT = 2 };

// One wavefront on one problem until the problem has no unclaimed rays and
// the wavefront's rays have ended (returns 0), or until the segment buffer
// nears its end and the rays are saved (returns 1). `resume` restores a saved
// wavefront. The SSP type is a template parameter, as in the single-problem
// kernel, so that each type runs specialised code. STAGE: where the problem's
// staged tables (SSP, boundaries, receiver depths) are read from -- 1, a copy
// in LDS that the compiler knows to be LDS (ds_read, which does not wait for
// the descriptor stores in flight); 2, global memory known to be global; 0,
// either, through a generic pointer (the fused mappings).
template<typename real, char ST, int MODE, int STAGE>
__device__ __attribute__((noinline)) int run_problem(const BatchDev<real> &Bdref, const DevOut &Oref, int pid,
                                                     const RaySave<real> *resume, SegBlock &blk_io, unsigned char *lds,
                                                     unsigned long long &all_steps, unsigned long long &all_contrib)
{
    // copies, for the same reason as the problem's below: through the
    // references their fields would be reloaded after every descriptor store;
    // the segment block in registers (through its reference, every append
    // would read it back from memory and wait for the stores in flight)
    const DevOut O = Oref;
    SegBlock blk = blk_io;
    const BatchDev<real> Bd = Bdref;
    const int lane = threadIdx.x;
    // a copy: through a reference the problem's fields would be reloaded after
    // every descriptor store (they might alias)
    Prob2D<real> P = Bd.probs[pid];
    const real *G = STAGE == 0 ? P.base : as_global(P.base);
    const real *B;
    if constexpr(STAGE == 1) {
        extern __shared__ __align__(16) unsigned char lds_dyn[];
        real *L = reinterpret_cast<real *>(lds_dyn);
        __syncthreads();
        for(int i = threadIdx.x; i < P.staged_reals; i += blockDim.x) L[i] = G[i];
        __syncthreads();
        B = L;
    } else if constexpr(STAGE == 2) {
        B = G;
    } else {
        B = stage_problem(P, Bd.lds_cap, lds);
    }
    if constexpr(STAGE != 0) { // (null stays null)
        P.ray_stats = as_global(P.ray_stats);
        P.ray_xend = as_global(P.ray_xend);
    }
    const int nr = P.njobs;
    bool pexh = false, active = false, init_fail = false;
    int result = 0;
    Err e = {0, 0, 0, 0};
    RayPt<real> p0;
    Trace<real> S;
    Infl<real> I;
    int job = 0, isz = 0, ia = 0, is = 0;
    unsigned long long steps = 0, contrib = 0;
    if(resume) {
        const RaySave<real> &sv = resume[lane];
        p0 = sv.p0;
        S = sv.S;
        I = sv.I;
        job = sv.job;
        is = sv.is;
        isz = sv.isz;
        ia = sv.ia;
        active = sv.active != 0;
    }
    for(int iter = 0;; ++iter) {
        if(!pexh) {
            unsigned long long need = __ballot(!active);
            if(need) {
                int cnt = __popcll(need);
                int leader = __ffsll((unsigned long long)need) - 1;
                int base = 0;
                if(lane == leader) base = atomicAdd(Bd.ray_next + pid, cnt);
                base = __shfl(base, leader);
                if(base + cnt >= nr) pexh = true;
                if(!active) {
                    int my = base + __popcll(need & ((1ull << lane) - 1ull));
                    if(my < nr) {
                        job = my;
                        JobIndices(job, P, isz, ia);
                        is = 0;
                        active = RayInit<real, ST>(isz, ia, p0, S, I, P, B, G, e);
                        init_fail = !active;
                        if(!active && P.ray_stats) {
                            P.ray_stats[3 * job] = 0;
                            P.ray_stats[3 * job + 1] = 0;
                            P.ray_stats[3 * job + 2] = 0;
                            P.ray_xend[2 * job] = 0.0;
                            P.ray_xend[2 * job + 1] = 0.0;
                        }
                    }
                }
                note_done(O.prog, init_fail);
                init_fail = false;
            }
        }
        if(!__any(active)) break;
        if((iter & 31) == 31) {
            // HasErrored of this problem (its rays stop, other problems go on),
            // and suspension when the segment buffer nears its end
            if(e.err) atomicOr(Bd.perr + 4 * pid, e.err);
            uint32_t ge = __hip_atomic_load(Bd.perr + 4 * pid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t sc = __hip_atomic_load(O.seg_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if(__any(ge != 0u || e.err != 0u)) {
                if(lane == 0) atomicMax(Bd.ray_next + pid, nr);
                break;
            }
            if(__builtin_amdgcn_readfirstlane(sc) >= Bd.suspend_at) {
                int u = 0;
                if(lane == 0) u = atomicAdd(Bd.nsave_out, 1);
                u = __shfl(u, 0);
                if(lane == 0) Bd.save_out_pid[u] = pid;
                RaySave<real> &sv = Bd.save_out[(size_t)u * WAVE + lane];
                sv.p0 = p0;
                sv.S = S;
                sv.I = I;
                sv.job = job;
                sv.is = is;
                sv.isz = isz;
                sv.ia = ia;
                sv.active = active ? 1 : 0;
                result = 1;
                break;
            }
        }
        RayPt<real> p1, p2;
        SegDesc<real> d;
        d.nitems = 0;
        bool two = false;
        if(active) {
            two = RayUpdate<real, ST>(p0, p1, p2, S, P, B, G, e);
            prelude<real, MODE != TRACE_DESC>(d, p0, p1, I, is, P, G, B + P.o_rzs, isz, ia);
            d.prob = pid;
            ++is;
        }
        if constexpr(MODE == TRACE_DESC) {
            append_seg(d, lane, O, blk, seg_full(P));
        } else {
            DevSink<real> sink{&O, &P, 0};
            seg_items<real, RT_TL, DevSink<real>, MODE == TRACE_FUSED_FAST>(d, P, G, P.Rz, sink);
            contrib += sink.nc;
        }
        if(__any(two)) {
            d.nitems = 0;
            if(active && two) {
                prelude<real, MODE != TRACE_DESC>(d, p1, p2, I, is, P, G, B + P.o_rzs, isz, ia);
                d.prob = pid;
                ++is;
            }
            if constexpr(MODE == TRACE_DESC) {
                append_seg(d, lane, O, blk, seg_full(P));
            } else {
                DevSink<real> sink{&O, &P, 0};
                seg_items<real, RT_TL, DevSink<real>, MODE == TRACE_FUSED_FAST>(d, P, G, P.Rz, sink);
                contrib += sink.nc;
            }
        }
        bool fin = false;
        if(active) {
            p0 = two ? p2 : p1;
            if(RayTerminate(p0, is, S, MAXN, P, e)) {
                active = false;
                fin = true;
                steps += (unsigned long long)is;
                if(P.ray_stats) {
                    P.ray_stats[3 * job] = is;
                    P.ray_stats[3 * job + 1] = p0.NumTopBnc;
                    P.ray_stats[3 * job + 2] = p0.NumBotBnc;
                    P.ray_xend[2 * job] = (double)p0.x.x;
                    P.ray_xend[2 * job + 1] = (double)p0.x.y;
                }
            }
        }
        note_done(O.prog, fin);
    }
    flush_problem_err(e, Bd.perr + 4 * pid);
    unsigned long long t = steps;
    for(int off = WAVE / 2; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if(lane == 0 && t) atomicAdd(Bd.psteps + pid, t);
    all_steps += steps;
    all_contrib += contrib;
    blk_io = blk;
    return result;
}

template<typename real, int MODE>
__device__ inline int run_problem_any(const BatchDev<real> &Bd, const DevOut &O, int pid, const RaySave<real> *resume,
                                      SegBlock &blk, unsigned char *lds, unsigned long long &all_steps,
                                      unsigned long long &all_contrib)
{
    const Prob2D<real> &Pp = Bd.probs[pid];
    const bool lds_fits = Pp.staged_reals > 0 && Pp.staged_reals <= Bd.lds_cap;
#define BHC_RUN(ST)                                                                                                   \
    if constexpr(MODE == TRACE_DESC) {                                                                                \
        if(lds_fits) return run_problem<real, ST, MODE, 1>(Bd, O, pid, resume, blk, lds, all_steps, all_contrib);    \
        return run_problem<real, ST, MODE, 2>(Bd, O, pid, resume, blk, lds, all_steps, all_contrib);                 \
    } else {                                                                                                          \
        return run_problem<real, ST, MODE, 0>(Bd, O, pid, resume, blk, lds, all_steps, all_contrib);                 \
    }
    switch(Pp.ssp_type) {
    case 'C': BHC_RUN('C')
    case 'P': BHC_RUN('P')
    case 'S': BHC_RUN('S')
    case 'N': BHC_RUN('N')
    case 'Q': BHC_RUN('Q')
    default: BHC_RUN('A')
    }
#undef BHC_RUN
}

template<typename real, int MODE>
__global__ void __launch_bounds__(64) k_trace_batch(BatchDev<real> Bd, DevOut O)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x;
    SegBlock blk{0u, 0u, 0u};
    unsigned long long all_steps = 0, all_contrib = 0;
    const int nsave_in = __builtin_amdgcn_readfirstlane(*Bd.nsave_in);
    for(;;) {
        // a saved wavefront of the previous pass first
        int u = -1;
        if(nsave_in > 0) {
            if(lane == 0) u = atomicAdd(Bd.resume_next, 1);
            u = __builtin_amdgcn_readfirstlane(__shfl(u, 0));
        }
        if(u >= 0 && u < nsave_in) {
            int pid = __builtin_amdgcn_readfirstlane(Bd.save_in_pid[u]);
            if(run_problem_any<real, MODE>(Bd, O, pid, Bd.save_in + (size_t)u * WAVE, blk, lds_raw, all_steps,
                                           all_contrib))
                break;
            continue;
        }
        // no new problems once the segment buffer nears its end: the next
        // pass continues from the work counters
        uint32_t sc = __hip_atomic_load(O.seg_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if(__builtin_amdgcn_readfirstlane(sc) >= Bd.suspend_at) {
            if(lane == 0) atomicOr(Bd.more, 1);
            break;
        }
        int p = 0;
        if(lane == 0) {
            p = __hip_atomic_load(Bd.pcursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while(p < Bd.nprob
                  && __hip_atomic_load(Bd.ray_next + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      >= Bd.probs[p].njobs) {
                atomicMax(Bd.pcursor, p + 1);
                ++p;
            }
        }
        p = __builtin_amdgcn_readfirstlane(__shfl(p, 0));
        if(p >= Bd.nprob) break;
        if(run_problem_any<real, MODE>(Bd, O, p, nullptr, blk, lds_raw, all_steps, all_contrib)) break;
    }
    seal_block<real>(blk, lane, O);
    for(int off = WAVE / 2; off > 0; off >>= 1) {
        all_steps += __shfl_xor(all_steps, off);
        all_contrib += __shfl_xor(all_contrib, off);
    }
    if(lane == 0) atomicAdd(O.counters + 0, all_steps);
    if(lane == 0 && all_contrib) atomicAdd(O.counters + 1, all_contrib);
}

// Scatter of a batched pass: one wavefront per segment, the segment's problem
// descriptor read through a wavefront-uniform pointer.
template<typename real, bool FAST>
__global__ void __launch_bounds__(256) k_scatter_batch(BatchDev<real> Bd, DevOut O)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    const int nwaves = gridDim.x * (blockDim.x / WAVE);
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    unsigned long long nc = 0;
    for(uint32_t s = (uint32_t)wave; s < n; s += (uint32_t)nwaves) {
        SegDesc<real> d;
        load_seg_core(O, s, d);
        if(d.nitems <= 0) continue;
        const Prob2D<real> P = Bd.probs[__builtin_amdgcn_readfirstlane(d.prob)];
        if(seg_full(P)) load_seg_rest(O, s, d);
        DevSink<real> sink{&O, &P, 0};
        scatter_seg<real, RT_TL, FAST>(d, P, sink, lane);
        nc += sink.nc;
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if(lane == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Scatter of a batched pass with one lane per segment: a wavefront loads 64
// descriptors at once (one memory round trip for all of them) and every lane
// walks its own segment's receivers. Batched sweeps of hat beams produce a
// segment per step with only a few receivers each, where a wavefront per
// segment would spend its time waiting for the descriptor. Four wavefronts
// per SIMD (128 VGPRs, a few spills in FP64): the lanes wait on descriptor,
// receiver and problem loads, and more wavefronts hide more of it (3% faster
// than three).
template<typename real, bool FAST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_scatter_lanes(BatchDev<real> Bd, DevOut O)
{
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    unsigned long long nc = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for(uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
        SegDesc<real> d;
        load_seg_core(O, s, d);
        const bool act = d.nitems > 0;
        const unsigned long long am = __ballot(act);
        if(!am) continue;
        // The segments of a wavefront mostly belong to one problem (slots are
        // reserved SEG_BLOCK at a time by one tracing wavefront): then its
        // descriptor is read once, with wavefront-uniform (scalar) loads,
        // instead of by every lane.
        const int pu = __builtin_amdgcn_readfirstlane(__shfl(d.prob, __ffsll(am) - 1));
        if(!__any(act && d.prob != pu)) {
            if(!act) continue;
            // a copy: read through a reference, the problem's fields would be
            // reloaded after every field atomic (they might alias)
            const Prob2D<real> P = Bd.probs[pu];
            if(seg_full(P)) load_seg_rest(O, s, d);
            DevSink<real> sink{&O, &P, 0};
            seg_items<real, RT_TL, DevSink<real>, FAST>(d, P, P.base, P.Rz, sink);
            nc += sink.nc;
        } else {
            if(!act) continue;
            const Prob2D<real> P = Bd.probs[d.prob];
            if(seg_full(P)) load_seg_rest(O, s, d);
            DevSink<real> sink{&O, &P, 0};
            seg_items<real, RT_TL, DevSink<real>, FAST>(d, P, P.base, P.Rz, sink);
            nc += sink.nc;
        }
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if((threadIdx.x & (WAVE - 1)) == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Scatter of a batched pass with the items of 64 segments flattened over the
// wavefront: the wavefront copies 64 consecutive descriptors to LDS with
// coalesced loads; the segments' range columns are spread over the lanes (a
// prefix sum of the column counts), each column's depth band is computed, and
// the receivers of all bands are spread over the lanes again, so every lane
// evaluates the same number of receivers whatever the segments' sizes.
// Ray-centred segments are walked by their own lane.
template<typename real, bool FAST>
__global__ void __launch_bounds__(128) k_scatter_flat(BatchDev<real> Bd, DevOut O)
{
    constexpr int NW = 2;                                  // wavefronts per workgroup
    constexpr int W16 = (int)(sizeof(Seg<real>) / 16);     // 16-byte words per descriptor
    __shared__ uint4 stage[NW][WAVE * W16];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    const Seg<real> *S = reinterpret_cast<const Seg<real> *>(stage[w]);
    unsigned long long nc = 0;
    const uint32_t gw = blockIdx.x * NW + (uint32_t)w, nw = gridDim.x * NW;
    for(uint32_t base = gw * WAVE; base < n; base += nw * WAVE) {
        const int cnt = (int)min((uint32_t)WAVE, n - base);
        // word k of the descriptors base .. base + cnt - 1 (consecutive lanes,
        // consecutive slots) into the stage's records
        for(int t = lane; t < cnt * SEG_NW<real>; t += WAVE) {
            const int k = t / cnt, i = t - k * cnt;
            stage[w][(size_t)i * W16 + k] = *reinterpret_cast<const uint4 *>(seg_word(O, k, base + i));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int ncol = 0;
        if(lane < cnt && S[lane].d.nitems > 0) {
            const SegDesc<real> &d = S[lane].d;
            const Prob2D<real> &P = Bd.probs[d.prob];
            if(P.isRayCen) {
                DevSink<real> sink{&O, &P, 0};
                seg_items<real, RT_TL, DevSink<real>, FAST>(d, P, P.base, P.Rz, sink);
                nc += sink.nc;
            } else {
                ncol = (d.ir1 - d.ir0) * d.dir + 1;
            }
        }
        int cexcl;
        const int ctot = wave_excl_scan(ncol, lane, cexcl);
        for(int c0 = 0; c0 < ctot; c0 += WAVE) {
            const int ct = c0 + lane;
            const int seg = wave_owner(cexcl, ct);
            const int sexcl = __shfl(cexcl, seg);
            int ir = 0, i0 = 0, i1 = -1;
            if(ct < ctot) {
                const SegDesc<real> &d = S[seg].d;
                const Prob2D<real> &P = Bd.probs[d.prob];
                cart_column(d, ct - sexcl, P, P.base, ir, i0, i1);
            }
            const int icnt = i1 >= i0 ? i1 - i0 + 1 : 0;
            int iexcl;
            const int itot = wave_excl_scan(icnt, lane, iexcl);
            for(int t0 = 0; t0 < itot; t0 += WAVE) {
                const int t = t0 + lane;
                const int own = wave_owner(iexcl, t);
                const int oir = __shfl(ir, own), oi0 = __shfl(i0, own), oex = __shfl(iexcl, own);
                const int oseg = __shfl(seg, own);
                if(t < itot) {
                    const SegDesc<real> &d = S[oseg].d;
                    const Prob2D<real> P = Bd.probs[d.prob];
                    DevSink<real> sink{&O, &P, 0};
                    item_cart_at<real, RT_TL, DevSink<real>, FAST>(d, oir, oi0 + (t - oex), P, P.base, P.Rz, sink);
                    nc += sink.nc;
                }
            }
        }
        // the stage is rewritten by the next group
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if(lane == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Per-pass counters reset by a kernel (the passes are replayed as HIP graphs,
// which take no memset nodes here).
__global__ void k_pass_reset(uint32_t *seg_count, uint32_t *overflow, int32_t *resume, int32_t *nsave_out,
                             int32_t *more)
{
    if(threadIdx.x == 0) {
        *seg_count = 0u;
        *overflow = 0u;
        *resume = 0;
        *nsave_out = 0;
        *more = 0;
    }
}

} // namespace

struct bhc_sweep {
    bhc_ctx *ctx;
    int precision, nprob;
    int with_stats;
    std::vector<int> njobs, nrz, nrr, nsz;
    std::vector<size_t> field_off, stats_off;   // in floats / rays
    size_t field_floats, total_rays;
    void *d_blobs;
    void *d_probs;
    float *d_rzsz;
    float *d_fields;
    int32_t *d_stats;
    double *d_xend;
    int32_t *d_ray_next, *d_pcursor, *d_resume, *d_nsave, *d_more;   // d_nsave: [2], one per parity
    uint32_t *d_perr;
    unsigned long long *d_psteps, *d_counters;
    void *d_save[2];
    int32_t *d_save_pid[2];
    int max_waves;
    int lds_bytes, lds_bytes_max;
    int passes;
    int scatter_blocks;
    int scatter_lanes;          // 0: one wavefront per segment, 1: one lane per segment, 2: 64 segments' receivers flattened over a wavefront
    double trace_ms, scatter_ms;
    // HIP graphs of one pass (reset, trace, scatter), one per parity of the
    // saved-ray buffers; rebuilt when the launch configuration changes
    bool use_graph;
    hipGraphExec_t graph[2];
    int graph_key[5];
    bool tuned;
};

extern "C" void bhc_sweep_release(bhc_sweep *sw)
{
    if(!sw) return;
    (void)hipSetDevice(sw->ctx->device);
    (void)hipStreamSynchronize(sw->ctx->stream);
    void *ptrs[] = {sw->d_blobs, sw->d_probs, sw->d_rzsz,  sw->d_fields,   sw->d_stats,  sw->d_xend,
                    sw->d_ray_next, sw->d_pcursor, sw->d_resume, sw->d_nsave, sw->d_more, sw->d_perr, sw->d_psteps,
                    sw->d_counters, sw->d_save[0], sw->d_save[1], sw->d_save_pid[0], sw->d_save_pid[1]};
    for(void *p : ptrs)
        if(p) (void)hipFree(p);
    for(int k = 0; k < 2; ++k)
        if(sw->graph[k]) (void)hipGraphExecDestroy(sw->graph[k]);
    delete sw;
}

namespace {

template<typename real>
int sweep_upload(bhc_sweep *sw, const bhc_problem2d *const *probs)
{
    bhc_ctx *ctx = sw->ctx;
    int n = sw->nprob;
    std::vector<Prob2D<real>> P(n);
    std::vector<real> blobs;
    std::vector<size_t> boff(n);
    std::vector<float> rzsz;
    std::vector<size_t> rzoff(n), szoff(n);
    size_t max_staged = 0;
    sw->field_off.resize(n);
    sw->stats_off.resize(n);
    sw->field_floats = 0;
    sw->total_rays = 0;
    for(int i = 0; i < n; ++i) {
        const bhc_problem2d *p = probs[i];
        std::vector<real> blob;
        std::string err;
        if(pack<real>(p, blob, P[i], err)) return fail("problem %d: %s", i, err.c_str());
        boff[i] = blobs.size();
        blobs.insert(blobs.end(), blob.begin(), blob.end());
        while(blobs.size() % 32) blobs.push_back((real)0);
        max_staged = std::max(max_staged, (size_t)P[i].staged_reals);
        rzoff[i] = rzsz.size();
        rzsz.insert(rzsz.end(), p->Rz, p->Rz + p->NRz);
        szoff[i] = rzsz.size();
        rzsz.insert(rzsz.end(), p->Sz, p->Sz + p->NSz);
        sw->njobs.push_back((p->iSingle_alpha >= 1 ? 1 : p->Nalpha) * p->NSz);
        sw->nsz.push_back(p->NSz);
        sw->nrz.push_back(p->NRz_per_range);
        sw->nrr.push_back(p->NRr);
        sw->field_off[i] = sw->field_floats;
        sw->field_floats += 2 * (size_t)p->NSz * p->NRz_per_range * p->NRr;
        sw->field_floats = (sw->field_floats + 63) / 64 * 64;
        sw->stats_off[i] = sw->total_rays;
        sw->total_rays += sw->njobs.back();
    }
    // LDS per workgroup: up to 40 KiB, so that four single-wavefront
    // workgroups fit a CU's 160 KiB; larger tables stay in global memory
    size_t cap_reals = std::min<size_t>((size_t)40960, (size_t)ctx->lds_max) / sizeof(real);
    size_t lds_reals = std::min(max_staged, cap_reals);
    sw->lds_bytes = (int)(lds_reals * sizeof(real));
    sw->lds_bytes_max = sw->lds_bytes;
    if(dalloc((real **)&sw->d_blobs, blobs.size()) || dalloc(&sw->d_rzsz, rzsz.size())
       || dalloc(&sw->d_fields, sw->field_floats) || dalloc((Prob2D<real> **)&sw->d_probs, (size_t)n)
       || dalloc(&sw->d_ray_next, (size_t)n) || dalloc(&sw->d_pcursor, 1) || dalloc(&sw->d_resume, 1)
       || dalloc(&sw->d_nsave, 2) || dalloc(&sw->d_more, 1) || dalloc(&sw->d_perr, 4 * (size_t)n)
       || dalloc(&sw->d_psteps, (size_t)n)
       || dalloc(&sw->d_counters, 2))
        return -1;
    if(sw->with_stats
       && (dalloc(&sw->d_stats, 3 * sw->total_rays) || dalloc(&sw->d_xend, 2 * sw->total_rays)))
        return -1;
    HIPCHECK(hipMemcpy(sw->d_blobs, blobs.data(), blobs.size() * sizeof(real), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(sw->d_rzsz, rzsz.data(), rzsz.size() * sizeof(float), hipMemcpyHostToDevice));
    for(int i = 0; i < n; ++i) {
        P[i].base = (const real *)sw->d_blobs + boff[i];
        if((size_t)P[i].staged_reals > lds_reals) P[i].staged_reals = 0;
        P[i].Rz = sw->d_rzsz + rzoff[i];
        P[i].Sz = sw->d_rzsz + szoff[i];
        P[i].field = sw->d_fields + sw->field_off[i];
        P[i].njobs = sw->njobs[i];
        if(sw->with_stats) {
            P[i].ray_stats = sw->d_stats + 3 * sw->stats_off[i];
            P[i].ray_xend = sw->d_xend + 2 * sw->stats_off[i];
        }
    }
    HIPCHECK(hipMemcpy(sw->d_probs, P.data(), n * sizeof(Prob2D<real>), hipMemcpyHostToDevice));
    // one single-wavefront workgroup per SIMD: the trace kernel runs at one
    // wavefront per SIMD (register bound), and 4 x 40 KiB of LDS fill a CU
    sw->max_waves = ctx->num_cus * 4;
    size_t save_units = (size_t)sw->max_waves;
    for(int k = 0; k < 2; ++k)
        if(dalloc((RaySave<real> **)&sw->d_save[k], save_units * WAVE) || dalloc(&sw->d_save_pid[k], save_units))
            return -1;
    sw->scatter_blocks = ctx->num_cus * 8;
    sw->scatter_lanes = 1;
    return 0;
}

// Device view of problems [p0, p1): problem indices inside the launch (work
// counters, saved rays, segment descriptors) are relative to p0.
template<typename real> BatchDev<real> batch_dev(bhc_sweep *sw, int p0, int p1)
{
    BatchDev<real> Bd;
    memset(&Bd, 0, sizeof(Bd));
    Bd.probs = (const Prob2D<real> *)sw->d_probs + p0;
    Bd.nprob = p1 - p0;
    Bd.lds_cap = sw->lds_bytes / (int)sizeof(real);
    Bd.ray_next = sw->d_ray_next + p0;
    Bd.pcursor = sw->d_pcursor;
    Bd.perr = sw->d_perr + 4 * (size_t)p0;
    Bd.psteps = sw->d_psteps + p0;
    Bd.resume_next = sw->d_resume;
    Bd.more = sw->d_more;
    return Bd;
}

template<typename real> DevOut batch_out(bhc_sweep *sw)
{
    bhc_ctx *c = sw->ctx;
    DevOut O;
    memset(&O, 0, sizeof(O));
    O.counters = sw->d_counters;
    O.segs = c->segs;
    O.seg_count = c->d_seg_count;
    O.overflow = c->d_overflow;
    O.seg_cap = (uint32_t)std::min<size_t>(c->seg_bytes / sizeof(Seg<real>), 0xFFFFFFF0u);
    O.prog = prog_of(c);
    return O;
}

// Enqueue one pass of parity `cur` (reads saved rays from buffer cur, saves
// into 1 - cur) on the context's stream.
template<typename real> void enqueue_pass(bhc_sweep *sw, BatchDev<real> Bd, const DevOut &O, int cur, hipEvent_t e_mid)
{
    bhc_ctx *c = sw->ctx;
    hipStream_t s = c->stream;
    Bd.save_in = (const RaySave<real> *)sw->d_save[cur];
    Bd.save_in_pid = sw->d_save_pid[cur];
    Bd.nsave_in = sw->d_nsave + cur;
    Bd.save_out = (RaySave<real> *)sw->d_save[1 - cur];
    Bd.save_out_pid = sw->d_save_pid[1 - cur];
    Bd.nsave_out = sw->d_nsave + (1 - cur);
    hipLaunchKernelGGL(k_pass_reset, dim3(1), dim3(64), 0, s, c->d_seg_count, c->d_overflow, sw->d_resume,
                       sw->d_nsave + (1 - cur), sw->d_more);
    if(sw->scatter_lanes == 3) {
        // fused: the trace evaluates every step's receivers itself
        if(c->exact_phasor)
            hipLaunchKernelGGL((k_trace_batch<real, TRACE_FUSED_EXACT>), dim3(sw->max_waves), dim3(WAVE), sw->lds_bytes,
                               s, Bd, O);
        else
            hipLaunchKernelGGL((k_trace_batch<real, TRACE_FUSED_FAST>), dim3(sw->max_waves), dim3(WAVE), sw->lds_bytes,
                               s, Bd, O);
        if(e_mid) (void)hipEventRecord(e_mid, s);
        return;
    }
    hipLaunchKernelGGL((k_trace_batch<real, TRACE_DESC>), dim3(sw->max_waves), dim3(WAVE), sw->lds_bytes, s, Bd, O);
    if(e_mid) (void)hipEventRecord(e_mid, s);
    if(sw->scatter_lanes == 2) {
        if(c->exact_phasor)
            hipLaunchKernelGGL((k_scatter_flat<real, false>), dim3(2 * sw->scatter_blocks), dim3(128), 0, s, Bd, O);
        else
            hipLaunchKernelGGL((k_scatter_flat<real, true>), dim3(2 * sw->scatter_blocks), dim3(128), 0, s, Bd, O);
    } else if(sw->scatter_lanes) {
        if(c->exact_phasor)
            hipLaunchKernelGGL((k_scatter_lanes<real, false>), dim3(sw->scatter_blocks), dim3(256), 0, s, Bd, O);
        else
            hipLaunchKernelGGL((k_scatter_lanes<real, true>), dim3(sw->scatter_blocks), dim3(256), 0, s, Bd, O);
    } else {
        if(c->exact_phasor)
            hipLaunchKernelGGL((k_scatter_batch<real, false>), dim3(sw->scatter_blocks), dim3(256), 0, s, Bd, O);
        else
            hipLaunchKernelGGL((k_scatter_batch<real, true>), dim3(sw->scatter_blocks), dim3(256), 0, s, Bd, O);
    }
}

// Trace + scatter problems [p0, p1) in one persistent launch (a sequence of
// passes): zero their fields and counters, then passes until no ray is left.
// Graph mode replays one captured graph per pass; eager mode times the trace
// and scatter kernels of every pass.
template<typename real> int sweep_core(bhc_sweep *sw, int p0, int p1, bool graph)
{
    bhc_ctx *c = sw->ctx;
    hipStream_t s = c->stream;
    if(ensure_segs(c)) return -1;
    BatchDev<real> Bd = batch_dev<real>(sw, p0, p1);
    DevOut O = batch_out<real>(sw);
    // every wavefront may add 2 segments per lane per step between two polls
    // (32 steps apart) plus one partly used slot block
    size_t margin = (size_t)sw->max_waves * (2 * 32 * WAVE + 2 * SEG_BLOCK);
    if(margin * 2 >= O.seg_cap) return fail("segment buffer too small for a batched sweep (raise BHC_SEG_BUFFER_MB)");
    Bd.suspend_at = (uint32_t)(O.seg_cap - margin);
    const int np = p1 - p0;
    {
        long long rays = 0;
        for(int i = p0; i < p1; ++i) rays += sw->njobs[i];
        if(prog_begin(c, rays)) return -1;
    }
    const size_t f0 = sw->field_off[p0], f1 = p1 == sw->nprob ? sw->field_floats : sw->field_off[p1];
    HIPCHECK(hipMemsetAsync(sw->d_fields + f0, 0, (f1 - f0) * sizeof(float), s));
    HIPCHECK(hipMemsetAsync(sw->d_ray_next + p0, 0, np * sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(sw->d_pcursor, 0, sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(sw->d_nsave, 0, 2 * sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(sw->d_perr + 4 * (size_t)p0, 0, 4 * np * sizeof(uint32_t), s));
    HIPCHECK(hipMemsetAsync(sw->d_psteps + p0, 0, np * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(sw->d_counters, 0, 2 * sizeof(unsigned long long), s));
    if(sw->with_stats) {
        const size_t r0 = sw->stats_off[p0], r1 = p1 == sw->nprob ? sw->total_rays : sw->stats_off[p1];
        HIPCHECK(hipMemsetAsync(sw->d_stats + 3 * r0, 0, 3 * (r1 - r0) * sizeof(int32_t), s));
        HIPCHECK(hipMemsetAsync(sw->d_xend + 2 * r0, 0, 2 * (r1 - r0) * sizeof(double), s));
    }
    if(graph) {
        int key[5] = {sw->scatter_blocks, sw->lds_bytes, p0, p1, sw->scatter_lanes};
        if(!sw->graph[0] || memcmp(key, sw->graph_key, sizeof(key)) != 0) {
            for(int k = 0; k < 2; ++k) {
                if(sw->graph[k]) (void)hipGraphExecDestroy(sw->graph[k]);
                sw->graph[k] = nullptr;
                hipGraph_t g = nullptr;
                HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                enqueue_pass<real>(sw, Bd, O, k, nullptr);
                HIPCHECK(hipStreamEndCapture(s, &g));
                hipError_t e = hipGraphInstantiate(&sw->graph[k], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHECK(e);
            }
            memcpy(sw->graph_key, key, sizeof(key));
        }
    }
    HIPCHECK(hipEventRecord(c->ev0, s));
    int cur = 0;
    sw->passes = 0;
    sw->trace_ms = sw->scatter_ms = 0.0;
    hipEvent_t e0, e1, e2;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventCreate(&e2));
    int rc = 0;
    for(;;) {
        (void)hipEventRecord(e0, s);
        if(graph) {
            if(hipGraphLaunch(sw->graph[cur], s) != hipSuccess) {
                rc = fail("sweep: graph launch failed");
                break;
            }
        } else {
            enqueue_pass<real>(sw, Bd, O, cur, e1);
        }
        (void)hipEventRecord(e2, s);
        if(hipGetLastError() != hipSuccess) {
            rc = fail("sweep: kernel launch failed");
            break;
        }
        int32_t ns = 0, more = 0;
        uint32_t of = 0;
        if(hipMemcpyAsync(&ns, sw->d_nsave + (1 - cur), sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess
           || hipMemcpyAsync(&more, sw->d_more, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess
           || hipMemcpyAsync(&of, c->d_overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess
           || hipStreamSynchronize(s) != hipSuccess) {
            rc = fail("sweep: pass failed");
            break;
        }
        float a = 0.f, b = 0.f;
        if(graph) {
            (void)hipEventElapsedTime(&a, e0, e2);
        } else {
            (void)hipEventElapsedTime(&a, e0, e1);
            (void)hipEventElapsedTime(&b, e1, e2);
        }
        sw->trace_ms += a;
        sw->scatter_ms += b;
        ++sw->passes;
        if(of) {
            rc = fail("sweep: segment buffer overflow (suspension margin too small)");
            break;
        }
        if(ns == 0 && !more) break;
        if(sw->passes > 100000) {
            rc = fail("sweep: no progress");
            break;
        }
        cur = 1 - cur;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    HIPCHECK(hipEventRecord(c->ev1, s));
    c->passes = sw->passes;
    if(!rc) c->prog_final = 1;
    return rc;
}

// Launch-configuration search on the sweep's first problems (eager passes,
// kernel times from HIP events): the scatter grid and whether the problems'
// tables are staged in LDS. Keeps the fastest.
template<typename real> int sweep_tune(bhc_sweep *sw)
{
    int nsub = std::min(sw->nprob, 64);
    int cus = sw->ctx->num_cus;
    const int grids[3] = {cus * 4, cus * 8, cus * 16};
    // staging budgets: the full one (one wavefront per SIMD), half of it (two
    // wavefronts per SIMD, problems with larger tables unstaged) and none
    const int rsz = (int)(sw->precision == 64 ? sizeof(double) : sizeof(float));
    const int half = (sw->lds_bytes_max / 2) / rsz * rsz;
    const int ldss[3] = {sw->lds_bytes_max, half, 0};
    auto timed = [&](int mode, int blocks, int lds, double &t) -> int {
        sw->scatter_lanes = mode;
        sw->scatter_blocks = blocks;
        sw->lds_bytes = lds;
        for(int rep = 0; rep < 2; ++rep) {
            if(sweep_core<real>(sw, 0, nsub, false)) return -1;
            if(rep == 1) t = sw->trace_ms + sw->scatter_ms;
        }
        return 0;
    };
    // the staging budget acts on the trace and the mapping / grid on the
    // scatter: choose the scatter's at the full budget, then the budget
    int best_blocks = grids[0], best_lanes = 1, best_lds = ldss[0];
    double best = 1e300, t = 0.0;
    for(int mode = 0; mode < 4; ++mode)
        for(int gi = 0; gi < (mode == 3 ? 1 : 3); ++gi) {   // the fused mode launches no scatter
            if(timed(mode, grids[gi], ldss[0], t)) return -1;
            if(t < best) {
                best = t;
                best_blocks = grids[gi];
                best_lanes = mode;
            }
        }
    for(int li = 1; li < (sw->lds_bytes_max > 0 ? 3 : 1); ++li) {
        if(timed(best_lanes, best_blocks, ldss[li], t)) return -1;
        if(t < best) {
            best = t;
            best_lds = ldss[li];
        }
    }
    sw->scatter_blocks = best_blocks;
    sw->lds_bytes = best_lds;
    sw->scatter_lanes = best_lanes;
    sw->tuned = true;
    return 0;
}

} // namespace

extern "C" int bhc_sweep_upload2d(bhc_ctx *ctx, const bhc_problem2d *const *probs, int nprob, int with_stats,
                                  bhc_sweep **out)
{
    *out = nullptr;
    if(!ctx || !probs || nprob < 1) return fail("bhc_sweep_upload2d: no problems");
    int prec = probs[0]->precision;
    for(int i = 0; i < nprob; ++i) {
        if(check_problem(probs[i])) return -1;
        char r0 = probs[i]->RunType[0];
        if(r0 != 'C' && r0 != 'S' && r0 != 'I') return fail("sweep problem %d: only TL run types are batched", i);
        if(probs[i]->precision != prec) return fail("sweep problem %d: all problems must share the precision", i);
    }
    HIPCHECK(hipSetDevice(ctx->device));
    bhc_sweep *sw = new bhc_sweep();
    sw->ctx = ctx;
    sw->precision = prec;
    sw->nprob = nprob;
    sw->with_stats = with_stats;
    sw->use_graph = true;
    if(const char *e = getenv("BHC_SWEEP_GRAPH")) sw->use_graph = atoi(e) != 0;
    int rc = prec == 64 ? sweep_upload<double>(sw, probs) : sweep_upload<float>(sw, probs);
    if(rc) {
        bhc_sweep_release(sw);
        return -1;
    }
    *out = sw;
    return 0;
}

extern "C" int bhc_sweep_launch(bhc_ctx *ctx, bhc_sweep *sw)
{
    if(check_idle(ctx, "bhc_sweep_launch")) return -1;
    if(!ctx || !sw || sw->ctx != ctx) return fail("bhc_sweep_launch: bad arguments");
    HIPCHECK(hipSetDevice(ctx->device));
    return sw->precision == 64 ? sweep_core<double>(sw, 0, sw->nprob, sw->use_graph)
                               : sweep_core<float>(sw, 0, sw->nprob, sw->use_graph);
}

extern "C" int bhc_sweep_launch_range(bhc_ctx *ctx, bhc_sweep *sw, int p0, int p1)
{
    if(check_idle(ctx, "bhc_sweep_launch_range")) return -1;
    if(!ctx || !sw || sw->ctx != ctx) return fail("bhc_sweep_launch_range: bad arguments");
    if(p0 < 0 || p1 > sw->nprob || p0 >= p1)
        return fail("bhc_sweep_launch_range: range [%d, %d) outside the sweep's %d problems", p0, p1, sw->nprob);
    HIPCHECK(hipSetDevice(ctx->device));
    return sw->precision == 64 ? sweep_core<double>(sw, p0, p1, sw->use_graph)
                               : sweep_core<float>(sw, p0, p1, sw->use_graph);
}

extern "C" int bhc_sweep_autotune(bhc_ctx *ctx, bhc_sweep *sw, int *scatter_blocks, int *lds_bytes,
                                  int *scatter_lanes)
{
    if(check_idle(ctx, "bhc_sweep_autotune")) return -1;
    if(!ctx || !sw || sw->ctx != ctx) return fail("bhc_sweep_autotune: bad arguments");
    HIPCHECK(hipSetDevice(ctx->device));
    int rc = sw->precision == 64 ? sweep_tune<double>(sw) : sweep_tune<float>(sw);
    if(scatter_blocks) *scatter_blocks = sw->scatter_blocks;
    if(lds_bytes) *lds_bytes = sw->lds_bytes;
    if(scatter_lanes) *scatter_lanes = sw->scatter_lanes;
    return rc;
}

extern "C" int bhc_sweep_set_launch(bhc_sweep *sw, int scatter_blocks, int lds_bytes)
{
    if(!sw) return fail("bhc_sweep_set_launch: null sweep");
    const int rsz = sw->precision == 64 ? (int)sizeof(double) : (int)sizeof(float);
    if(scatter_blocks < 1 || scatter_blocks > (1 << 20)) return fail("bhc_sweep_set_launch: scatter_blocks out of range");
    if(lds_bytes < 0 || lds_bytes > sw->lds_bytes_max || lds_bytes % rsz)
        return fail("bhc_sweep_set_launch: lds_bytes must be a multiple of %d in [0, %d]", rsz, sw->lds_bytes_max);
    sw->scatter_blocks = scatter_blocks;
    sw->lds_bytes = lds_bytes;
    sw->tuned = true;
    return 0;
}

extern "C" int bhc_sweep_set_scatter(bhc_sweep *sw, int mapping)
{
    if(!sw || mapping < 0 || mapping > 3) return fail("bhc_sweep_set_scatter: bad arguments");
    sw->scatter_lanes = mapping;
    return 0;
}

extern "C" int bhc_sweep_set_graph(bhc_sweep *sw, int use_graph)
{
    if(!sw) return fail("bhc_sweep_set_graph: null sweep");
    sw->use_graph = use_graph != 0;
    return 0;
}

extern "C" int bhc_sweep_stats(bhc_sweep *sw, double *trace_ms, double *scatter_ms, int *passes, long long *steps,
                               long long *contrib)
{
    if(!sw) return fail("bhc_sweep_stats: null sweep");
    unsigned long long cnt[2] = {0, 0};
    HIPCHECK(hipStreamSynchronize(sw->ctx->stream));
    HIPCHECK(hipMemcpy(cnt, sw->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if(trace_ms) *trace_ms = sw->trace_ms;
    if(scatter_ms) *scatter_ms = sw->scatter_ms;
    if(passes) *passes = sw->passes;
    if(steps) *steps = (long long)cnt[0];
    if(contrib) *contrib = (long long)cnt[1];
    return 0;
}

extern "C" int bhc_sweep_download(bhc_ctx *ctx, bhc_sweep *sw, int iprob, bhc_field_out *out)
{
    if(!ctx || !sw || iprob < 0 || iprob >= sw->nprob || !out) return fail("bhc_sweep_download: bad arguments");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    size_t nf = 2 * (size_t)sw->nsz[iprob] * sw->nrz[iprob] * sw->nrr[iprob];
    if(out->field)
        HIPCHECK(hipMemcpy(out->field, sw->d_fields + sw->field_off[iprob], nf * sizeof(float),
                           hipMemcpyDeviceToHost));
    uint32_t ew[4];
    HIPCHECK(hipMemcpy(ew, sw->d_perr + 4 * iprob, sizeof(ew), hipMemcpyDeviceToHost));
    out->err = ew[0];
    out->warn = ew[1];
    out->err_count = ew[2];
    out->warn_count = ew[3];
    unsigned long long st = 0;
    HIPCHECK(hipMemcpy(&st, sw->d_psteps + iprob, sizeof(st), hipMemcpyDeviceToHost));
    out->total_steps = (int64_t)st;
    out->total_contrib = 0;
    if(out->ray_nsteps && sw->with_stats) {
        int nj = sw->njobs[iprob];
        std::vector<int32_t> stv(3 * (size_t)nj);
        HIPCHECK(hipMemcpy(stv.data(), sw->d_stats + 3 * sw->stats_off[iprob], stv.size() * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
        for(int j = 0; j < nj; ++j) {
            out->ray_nsteps[j] = stv[3 * j];
            if(out->ray_ntop) out->ray_ntop[j] = stv[3 * j + 1];
            if(out->ray_nbot) out->ray_nbot[j] = stv[3 * j + 2];
        }
        if(out->ray_xend)
            HIPCHECK(hipMemcpy(out->ray_xend, sw->d_xend + 2 * sw->stats_off[iprob], 2 * (size_t)nj * sizeof(double),
                               hipMemcpyDeviceToHost));
    }
    return 0;
}

#include "field3d.h"
