// field3d.h -- kernels and launch layer of the full-3D TL path. Included at
// the end of trace2d.hip: it shares the device context, the segment buffer
// and the segment-block allocator of the 2D path.
//
// Same split as in 2D: k_trace3d runs one ray per lane (persistent lanes
// refilled from a device work counter) and reduces every step to a Seg3D
// descriptor; k_scatter3d gives each segment one wavefront and flattens its
// (range, bearing) rows' depth windows over the 64 lanes. Replaces
// RunFieldModesImpl<CFG, true, true> (src/mode/fieldimpl.hpp:27-28).
#include "core3d.h"
#include "pack3d.h"

namespace {

using namespace bhc3d;

template<typename real> struct alignas(16) Seg3 {
    Seg3D<real> d;
};
template<typename real> struct alignas(16) Seg3RW {
    Seg3R<real> d;
};
// Stored descriptor type of a 3D segment: Cartesian or ray-centred.
template<typename real, bool RCEN> using SegSlot3 = std::conditional_t<RCEN, Seg3RW<real>, Seg3<real>>;

struct DevOut3 {
    int32_t *stats;             // [njobs][3] steps, top / bottom bounces (or null)
    double *xend;               // [njobs][3]
    unsigned long long *counters; // [0] steps, [1] contributions
    uint32_t *errw;
    int32_t *work;
    int32_t job_begin, job_end;
    void *segs;
    uint32_t *seg_count;
    uint32_t seg_cap;
    uint32_t *overflow;
    ArrCand *cand;              // arrivals: [receivers][cand_cap] raw AddArr calls
    int32_t *ncand;
    int32_t cand_cap;
    int32_t eigen_memsize;
    bhc_eigen_hit *hits;        // eigen hits
    int32_t *neigen;
    Progress prog;              // finished-ray progress (as in the 2D trace)
};

struct Sink3 {
    float *field;
    unsigned long long nc;
    const DevOut3 *O;
    __device__ void add3(size_t idx, float re, float im)
    {
        unsafeAtomicAdd(field + 2 * idx, re);
        unsafeAtomicAdd(field + 2 * idx + 1, im);
    }
    __device__ void count() { ++nc; }
    __device__ void mark() {}
    // AddArr (arrivals.hpp:45-130) as a candidate keyed by (ray, step); the
    // merge kernel replays the single-threaded rule per receiver. Angles as
    // RayInit (trace.hpp:99) and ReceiverAngles (influence.hpp:123-136, 3D).
    template<typename real>
    __device__ void arr3(size_t idx, real amp, real phase, cpx<real> delay, real decl, real azim, int job, int is,
                         int ntop, int nbot, const Prob3D<real> &P)
    {
        int slot = atomicAdd(O->ncand + idx, 1);
        if(slot >= O->cand_cap) return;
        int ia, ib, isy, isx, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        const real RadDeg = RL(180.0) / RC<real>::PI;
        ArrCand c;
        c.key = ((long long)job << 17) | (long long)is;
        c.a = (double)amp;
        c.phase = (double)phase;
        c.dre = (double)delay.re;
        c.dim = (double)delay.im;
        c.src = (float)(RadDeg * P.base[P.o_alpha + ia]);
        c.srcaz = (float)(RadDeg * P.base[P.o_beta + ib]);
        c.rcv = (float)decl;
        c.rcvaz = (float)azim;
        c.ntop = ntop;
        c.nbot = nbot;
        O->cand[idx * (size_t)O->cand_cap + slot] = c;
    }
    // RecordEigenHit, eigenrays.hpp:24-41
    template<typename real>
    __device__ void eig3(int itheta, int ir, int iz, int job, int is, const Prob3D<real> &P)
    {
        int mi = atomicAdd(O->neigen, 1);
        if(mi >= O->eigen_memsize) return;
        int ia, ib, isy, isx, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        bhc_eigen_hit h;
        h.is = is;
        h.itheta = itheta;
        h.ir = ir;
        h.iz = iz;
        h.isx = isx;
        h.isy = isy;
        h.isz = isz;
        h.ialpha = ia;
        h.ibeta = ib;
        O->hits[mi] = h;
    }
};

template<typename real, bool RCEN = false> __device__ inline void seal3(const SegBlock &B, int lane, const DevOut3 &O)
{
    SegSlot3<real, RCEN> *segs = reinterpret_cast<SegSlot3<real, RCEN> *>(O.segs);
    for(uint32_t k = B.used + (uint32_t)lane; k < B.cap; k += WAVE) {
        uint32_t slot = B.base + k;
        if(slot < O.seg_cap) segs[slot].d.nitems = 0;
    }
}

template<typename real, typename Desc, bool RCEN = false>
__device__ inline void append3(const Desc &d, int lane, const DevOut3 &O, SegBlock &B)
{
    bool has = d.nitems > 0;
    unsigned long long m = __ballot(has);
    if(!m) return;
    uint32_t cnt = (uint32_t)__popcll(m);
    if(B.used + cnt > B.cap) {
        seal3<real, RCEN>(B, lane, O);
        uint32_t base = 0;
        if(lane == 0) base = atomicAdd(O.seg_count, SEG_BLOCK);
        B.base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        B.used = 0;
        B.cap = SEG_BLOCK;
    }
    if(has) {
        uint32_t slot = B.base + B.used + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if(slot < O.seg_cap) reinterpret_cast<SegSlot3<real, RCEN> *>(O.segs)[slot].d = d;
        else atomicOr(O.overflow, 1u);
    }
    B.used += cnt;
}

// Trace kernel: the ray loop of MainFieldModes (src/trace.hpp:525-589), 3D.
// Reflection tables staged in LDS up to this many reals (top and bottom together).
constexpr int REFL_LDS = 768;

template<typename real, bool RCEN = false> __global__ void __launch_bounds__(64) k_trace3d(Prob3D<real> P, DevOut3 O)
{
    __shared__ real refl_lds[REFL_LDS];
    const real *tabs = stage_refl_tables(P, refl_lds, REFL_LDS);
    const int lane = threadIdx.x & (WAVE - 1);
    const int njobs = O.job_end - O.job_begin;
    Err e = {0, 0, 0, 0};
    Ray3<real> p0;
    Trace3<real> S;
    Infl3<real> I;
    int job = 0, ia = 0, ib = 0, isx = 0, isy = 0, isz = 0, is = 0;
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
                        JobIndices3(job, P, ia, ib, isy, isx, isz);
                        is = 0;
                        active = RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e);
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
        Ray3<real> p1, p2;
        std::conditional_t<RCEN, Seg3R<real>, Seg3D<real>> d;
        d.nitems = 0;
        bool two = false;
        if(active) {
            two = RayUpdate3(p0, p1, p2, S, P, e, tabs);
            if constexpr(RCEN) prelude3rc(d, p0, p1, I, is, P, isz, isx, isy);
            else prelude3(d, p0, p1, I, is, P, isz, isx, isy);
            d.job = job;
            ++is;
        }
        append3<real, decltype(d), RCEN>(d, lane, O, blk);
        if(__any(two)) {
            d.nitems = 0;
            if(active && two) {
                if constexpr(RCEN) prelude3rc(d, p1, p2, I, is, P, isz, isx, isy);
                else prelude3(d, p1, p2, I, is, P, isz, isx, isy);
                d.job = job;
                ++is;
            }
            append3<real, decltype(d), RCEN>(d, lane, O, blk);
        }
        bool fin = false;
        if(active) {
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e)) {
                active = false;
                fin = true;
                steps += (unsigned long long)is;
                if(O.stats) {
                    O.stats[3 * job] = is;
                    O.stats[3 * job + 1] = p0.NumTopBnc;
                    O.stats[3 * job + 2] = p0.NumBotBnc;
                    O.xend[3 * job] = (double)p0.x.x;
                    O.xend[3 * job + 1] = (double)p0.x.y;
                    O.xend[3 * job + 2] = (double)p0.x.z;
                }
            }
        }
        note_done(O.prog, fin);
    }
    seal3<real, RCEN>(blk, lane, O);
    if(e.err) atomicOr(O.errw + 0, e.err);
    if(e.warn) atomicOr(O.errw + 1, e.warn);
    if(e.nerr) atomicAdd(O.errw + 2, e.nerr);
    if(e.nwarn) atomicAdd(O.errw + 3, e.nwarn);
    for(int off = WAVE / 2; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
    if(lane == 0) atomicAdd(O.counters + 0, steps);
}

// Scatter kernel: one wavefront per segment. Rows are the (range, bearing)
// pairs the segment's receiver walk visits (influence.hpp:1298-1386, 3D);
// each row's depth window is flattened over the lanes.
// STAGED: the receiver ranges, bearing table and depths are copied to LDS (the
// launch checks they fit) and read only from there -- compile-time, so the
// reads are LDS reads rather than generic ones. Rows take the circle band
// only: with a wavefront per segment the per-axis bounds cost more than the
// receivers they remove (260 against 139 ms per shelf-canyon pass). Three wavefronts per SIMD
// (at most 168 VGPRs): the per-segment chain of loads and FP64 row bounds is
// latency-bound, and the third wavefront hides more of it (14 % faster than
// two on the shelf-canyon field; four spill and are slower).
template<typename real, bool FAST, int RT = RT_TL, bool STAGED = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_scatter3d(Prob3D<real> P, DevOut3 O)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    const int nwaves = gridDim.x * (blockDim.x / WAVE);
    extern __shared__ __align__(16) unsigned char lds3[];
    const real *Rr, *trcv;
    const float *Rz;
    if constexpr(STAGED) {
        real *L = reinterpret_cast<real *>(lds3);
        float *Lz = reinterpret_cast<float *>(L + P.NRr + 2 * P.Ntheta);
        for(int i = threadIdx.x; i < P.NRr; i += blockDim.x) L[i] = P.base[P.o_Rr + i];
        for(int i = threadIdx.x; i < 2 * P.Ntheta; i += blockDim.x) L[P.NRr + i] = P.base[P.o_trcv + i];
        for(int i = threadIdx.x; i < P.NRz; i += blockDim.x) Lz[i] = P.Rz[i];
        __syncthreads();
        Rr = L;
        trcv = L + P.NRr;
        Rz = Lz;
    } else {
        Rr = P.base + P.o_Rr;
        trcv = P.base + P.o_trcv;
        Rz = P.Rz;
    }
    __shared__ int cand_all[4][WAVE];   // candidate bearings of each wavefront's segment
    int *cand = cand_all[(threadIdx.x / WAVE) & 3];
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    Sink3 sink{P.field, 0, &O};
    const Seg3<real> *segs = reinterpret_cast<const Seg3<real> *>(O.segs);
    for(uint32_t si = (uint32_t)wave; si < n; si += (uint32_t)nwaves) {
        const Seg3D<real> d = segs[si].d;
        if(d.nitems <= 0) continue;
        const int nr = (d.ir1 - d.ir0) * d.dir + 1;
        // bearings that can hold an accepted receiver (all of them beyond 64)
        int nb = P.Ntheta;
        const bool pre = P.Ntheta <= WAVE;
        if(pre) {
            real r0 = Rr[d.ir0], r1 = Rr[d.ir1];
            real rlo = rmin(r0, r1), rhi = rmax(r0, r1);
            real zs = P.rz_absmax + std::fabs(d.xa.z);
            bool c = lane < P.Ntheta && bearing3(d, trcv[2 * lane], trcv[2 * lane + 1], rlo, rhi, zs, P);
            unsigned long long m = __ballot(c);
            if(c) cand[__popcll(m & ((1ull << lane) - 1ull))] = lane;
            nb = __popcll(m);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const int nrow = nr * nb;
        for(int g = 0; g < nrow; g += WAVE) {
            const int row = g + lane;
            int ir = 0, it = 0, i0 = 0, i1 = -1;
            real xr = 0, yr = 0, zmin = 0, zmax = 0;
            if(row < nrow) {
                int j = row / nb;
                int k = row - j * nb;
                it = pre ? cand[k] : k;
                ir = d.ir0 + j * d.dir;
                real wlo, whi;
                if(row3(d, ir, it, P, Rr, trcv, xr, yr, zmin, zmax) && band3_circle(d, xr, yr, zmin, zmax, P, wlo, whi)) {
                    if(P.rz_sorted && wlo == wlo && whi == whi) {
                        depth_window(Rz, P.NRz, P.rz_rdz, wlo, whi, i0, i1);
                    } else {
                        i0 = 0;
                        i1 = P.NRz - 1;
                    }
                }
            }
            const int cnt = i1 >= i0 ? i1 - i0 + 1 : 0;
            int excl;
            const int total = wave_excl_scan(cnt, lane, excl);
            for(int t0 = 0; t0 < total; t0 += WAVE) {
                const int t = t0 + lane;
                const int own = wave_owner(excl, t);
                const int oir = __shfl(ir, own), oit = __shfl(it, own), oi0 = __shfl(i0, own);
                const int oexcl = __shfl(excl, own);
                const real oxr = __shfl(xr, own), oyr = __shfl(yr, own);
                const real ozmin = __shfl(zmin, own), ozmax = __shfl(zmax, own);
                if(t < total)
                    item3<real, Sink3, FAST, RT>(d, oir, oit, oi0 + (t - oexcl), oxr, oyr, ozmin, ozmax, P, Rz, sink);
            }
        }
    }
    unsigned long long nc = sink.nc;
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if(lane == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Index of the k-th set bit of m (k < popcount(m)), by halving.
__device__ inline int nth_set_bit(unsigned long long m, int k)
{
    int pos = 0;
    for(int w = 32; w >= 1; w >>= 1) {
        const unsigned long long low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if(k >= c) {
            k -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// Makes one wavefront's LDS writes visible to its other lanes (and orders its
// earlier LDS reads before later writes).
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Segments per wavefront group in k_scatter3d_flat.
constexpr int SCAT3_G = 16;

// Scatter with SCAT3_G consecutive segments per wavefront: their descriptors
// are copied to LDS by coalesced loads; every (segment, bearing) pair goes
// through the bearing prefilter, one pair per lane; the surviving (segment,
// range, bearing) rows of all the group's segments are spread over the lanes
// (a prefix sum of the row counts), each row's tests and depth window are
// computed, and the receivers of all rows are spread over the lanes again.
// The tests and contributions are those of k_scatter3d; the lanes stay busy
// when each segment has only a few rows. The receiver ranges, bearing table
// and depths are staged in LDS (the launch checks they fit). Four wavefronts
// per SIMD (at most 128 VGPRs): 4 % faster than three on the shelf-canyon
// field; 32 segments per group instead of 16 halve the occupancy through LDS
// and are 22 % slower.
template<typename real, bool FAST, int RT = RT_TL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_scatter3d_flat(Prob3D<real> P, DevOut3 O)
{
    constexpr int G = SCAT3_G;
    constexpr int W16 = (int)(sizeof(Seg3<real>) / 16);   // 16-byte words per descriptor
    __shared__ uint4 stage_all[4][G * W16];
    __shared__ unsigned long long bmask_all[4][G];
    __shared__ int2 rowq_all[4][2 * WAVE];
    __shared__ int2 accq_all[4][2 * WAVE];
    extern __shared__ __align__(16) unsigned char lds3f[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wib = threadIdx.x / WAVE;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / WAVE) + wib));
    const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
    real *L = reinterpret_cast<real *>(lds3f);
    float *Lz = reinterpret_cast<float *>(L + P.NRr + 2 * P.Ntheta);
    for(int i = threadIdx.x; i < P.NRr; i += blockDim.x) L[i] = P.base[P.o_Rr + i];
    for(int i = threadIdx.x; i < 2 * P.Ntheta; i += blockDim.x) L[P.NRr + i] = P.base[P.o_trcv + i];
    for(int i = threadIdx.x; i < P.NRz; i += blockDim.x) Lz[i] = P.Rz[i];
    __syncthreads();
    const real *Rr = L, *trcv = L + P.NRr;
    const float *Rz = Lz;
    uint4 *stage = stage_all[wib];
    unsigned long long *bmask = bmask_all[wib];
    int2 *rowq = rowq_all[wib];
    int2 *accq = accq_all[wib];
    const Seg3<real> *S = reinterpret_cast<const Seg3<real> *>(stage);
    const uint4 *gw = reinterpret_cast<const uint4 *>(O.segs);
    const bool pre = P.Ntheta <= WAVE;
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    Sink3 sink{P.field, 0, &O};
    for(uint32_t base = wave * G; base < n; base += nwaves * G) {
        const int ng = (int)min((uint32_t)G, n - base);
        for(int t = lane; t < ng * W16; t += WAVE) stage[t] = gw[(size_t)base * W16 + t];
        if(lane < G) bmask[lane] = 0ull;
        wave_lds_sync();
        if(pre) {
            // bearing prefilter, one (segment, bearing) pair per lane
            const int npair = ng * P.Ntheta;
            for(int p0 = 0; p0 < npair; p0 += WAVE) {
                const int p = p0 + lane;
                if(p < npair) {
                    const int g = p / P.Ntheta, t = p - g * P.Ntheta;
                    const Seg3D<real> &d = S[g].d;
                    if(d.nitems > 0) {
                        const real r0 = Rr[d.ir0], r1 = Rr[d.ir1];
                        const real zs = P.rz_absmax + std::fabs(d.xa.z);
                        if(bearing3(d, trcv[2 * t], trcv[2 * t + 1], rmin(r0, r1), rmax(r0, r1), zs, P))
                            atomicOr(bmask + g, 1ull << t);
                    }
                }
            }
            wave_lds_sync();
        }
        int nrow = 0, nb = 0;
        unsigned long long bm = 0ull;
        if(lane < ng && S[lane].d.nitems > 0) {
            const Seg3D<real> &d = S[lane].d;
            bm = bmask[lane];
            nb = pre ? __popcll(bm) : P.Ntheta;
            nrow = ((d.ir1 - d.ir0) * d.dir + 1) * nb;
        }
        // receivers that pass the beam tests are queued in LDS as
        // (iz NRr + ir, g Ntheta + it) and their contributions made 64 at a
        // time (item3 again, with its row recomputed), so the contribution
        // work runs on full wavefronts
        int qa = 0;
        auto do_acc = [&](int k) {
            if(lane < k) {
                const int2 e = accq[lane];
                const int iz = e.x / P.NRr, ir = e.x - iz * P.NRr;
                const int g = e.y / P.Ntheta, it = e.y - g * P.Ntheta;
                const Seg3D<real> &d = S[g].d;
                real xr, yr, zmin, zmax;
                row3(d, ir, it, P, Rr, trcv, xr, yr, zmin, zmax);
                item3<real, Sink3, FAST, RT>(d, ir, it, iz, xr, yr, zmin, zmax, P, Rz, sink);
            }
        };
        // rows that pass the row test are queued in LDS as (ir, g Ntheta + it)
        // and taken 64 at a time, so the band and depth-window work runs on
        // full wavefronts
        auto do_rows = [&](int k) {
            int g = 0, ir = 0, it = 0, i0 = 0, i1 = -1;
            real xr = 0, yr = 0, zmin = 0, zmax = 0;
            if(lane < k) {
                const int2 e = rowq[lane];
                ir = e.x;
                g = e.y / P.Ntheta;
                it = e.y - g * P.Ntheta;
                const Seg3D<real> &d = S[g].d;
                real wlo, whi;
                if(row3(d, ir, it, P, Rr, trcv, xr, yr, zmin, zmax) && band3(d, xr, yr, zmin, zmax, P, wlo, whi)) {
                    if(P.rz_sorted && wlo == wlo && whi == whi) {
                        depth_window(Rz, P.NRz, P.rz_rdz, wlo, whi, i0, i1);
                    } else {
                        i0 = 0;
                        i1 = P.NRz - 1;
                    }
                }
            }
            const int cnt = i1 >= i0 ? i1 - i0 + 1 : 0;
            int excl;
            const int total = wave_excl_scan(cnt, lane, excl);
            for(int t0 = 0; t0 < total; t0 += WAVE) {
                const int t = t0 + lane;
                const int own = wave_owner(excl, t);
                const int og = __shfl(g, own), oir = __shfl(ir, own), oit = __shfl(it, own);
                const int oi0 = __shfl(i0, own), oexcl = __shfl(excl, own);
                const real oxr = __shfl(xr, own), oyr = __shfl(yr, own);
                const real ozmin = __shfl(zmin, own), ozmax = __shfl(zmax, own);
                const int oiz = oi0 + (t - oexcl);
                const bool acc = t < total && item3_accepts(S[og].d, oxr, oyr, oiz, ozmin, ozmax, P, Rz);
                const unsigned long long am = __ballot(acc);
                if(acc) accq[qa + __popcll(am & ((1ull << lane) - 1ull))] = make_int2(oiz * P.NRr + oir, og * P.Ntheta + oit);
                qa += __popcll(am);
                if(qa >= WAVE) {
                    wave_lds_sync();
                    do_acc(WAVE);
                    int2 v = make_int2(0, 0);
                    const bool mv = lane < qa - WAVE;
                    if(mv) v = accq[WAVE + lane];
                    wave_lds_sync();
                    if(mv) accq[lane] = v;
                    qa -= WAVE;
                    wave_lds_sync();
                }
            }
        };
        int rexcl;
        const int nrows = wave_excl_scan(nrow, lane, rexcl);
        int qn = 0;
        for(int r0 = 0; r0 < nrows; r0 += WAVE) {
            const int rt = r0 + lane;
            const int g = wave_owner(rexcl, rt);
            const int loc = rt - __shfl(rexcl, g);
            const int gnb = __shfl(nb, g);
            const unsigned long long gbm = __shfl(bm, g);
            bool pass = false;
            int ir = 0, it = 0;
            if(rt < nrows) {
                const Seg3D<real> &d = S[g].d;
                const int j = loc / gnb, k = loc - j * gnb;
                it = pre ? nth_set_bit(gbm, k) : k;
                ir = d.ir0 + j * d.dir;
                real xr, yr, zmin, zmax;
                pass = row3(d, ir, it, P, Rr, trcv, xr, yr, zmin, zmax);
            }
            const unsigned long long pm = __ballot(pass);
            if(pass) rowq[qn + __popcll(pm & ((1ull << lane) - 1ull))] = make_int2(ir, g * P.Ntheta + it);
            qn += __popcll(pm);
            if(qn >= WAVE) {
                wave_lds_sync();
                do_rows(WAVE);
                int2 v = make_int2(0, 0);
                const bool mv = lane < qn - WAVE;
                if(mv) v = rowq[WAVE + lane];
                wave_lds_sync();
                if(mv) rowq[lane] = v;
                qn -= WAVE;
                wave_lds_sync();
            }
        }
        if(qn > 0) {
            wave_lds_sync();
            do_rows(qn);
        }
        if(qa > 0) {
            wave_lds_sync();
            do_acc(qa);
        }
        // the stage is rewritten by the next group
        wave_lds_sync();
    }
    unsigned long long nc = sink.nc;
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if(lane == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Scatter of ray-centred 3D segments (Step_InfluenceGeoRayCen,
// influence.hpp:1030-1194): one wavefront per segment; the (depth, bearing)
// rows are spread over the lanes, each lane finds its row's bracketed
// receiver ranges from the ray normals at both ends, and the receivers of all
// rows are flattened over the lanes by a prefix sum.
template<typename real, int RT>
__global__ void __launch_bounds__(256) k_scatter3rc(Prob3D<real> P, DevOut3 O)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    const int nwaves = gridDim.x * (blockDim.x / WAVE);
    const real *Rr = P.base + P.o_Rr, *trcv = P.base + P.o_trcv;
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    Sink3 sink{P.field, 0, &O};
    const Seg3RW<real> *segs = reinterpret_cast<const Seg3RW<real> *>(O.segs);
    const int nrow = P.NRz * P.Ntheta;
    for(uint32_t si = (uint32_t)wave; si < n; si += (uint32_t)nwaves) {
        const Seg3R<real> d = segs[si].d;
        if(d.nitems <= 0) continue;
        for(int g = 0; g < nrow; g += WAVE) {
            const int row = g + lane;
            int iz = 0, it = 0, first = 0, step = 1, cnt = 0;
            real mA = 0, nA = 0, rA = 0, mB = 0, nB = 0, rB = 0;
            if(row < nrow) {
                iz = row / P.Ntheta;
                it = row - iz * P.Ntheta;
                cnt = row3rc(d, iz, it, P, Rr, trcv, P.Rz, mA, nA, rA, mB, nB, rB, first, step);
            }
            int excl;
            const int total = wave_excl_scan(cnt, lane, excl);
            for(int t0 = 0; t0 < total; t0 += WAVE) {
                const int t = t0 + lane;
                const int own = wave_owner(excl, t);
                const int oiz = __shfl(iz, own), oit = __shfl(it, own), ofirst = __shfl(first, own);
                const int ostep = __shfl(step, own), oex = __shfl(excl, own);
                const real omA = __shfl(mA, own), onA = __shfl(nA, own), orA = __shfl(rA, own);
                const real omB = __shfl(mB, own), onB = __shfl(nB, own), orB = __shfl(rB, own);
                if(t < total)
                    item3rc<real, Sink3, false, RT>(d, oiz, oit, ofirst + (t - oex) * ostep, omA, onA, orA, omB, onB,
                                                    orB, P, Rr, sink);
            }
        }
    }
    unsigned long long nc = sink.nc;
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if(lane == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Scatter with one lane per segment: each lane walks its own segment's rows
// and depth windows serially (no cross-lane flattening), with the circle band
// (the per-axis bounds: 199 against 156 ms per shelf-canyon pass).
template<typename real, bool FAST, int RT = RT_TL>
__global__ void __launch_bounds__(256) k_scatter3d_lanes(Prob3D<real> P, DevOut3 O)
{
    extern __shared__ __align__(16) unsigned char lds3l[];
    const real *Rr = P.base + P.o_Rr, *trcv = P.base + P.o_trcv;
    const float *Rz = P.Rz;
    if(P.stage_tables) {
        real *L = reinterpret_cast<real *>(lds3l);
        float *Lz = reinterpret_cast<float *>(L + P.NRr + 2 * P.Ntheta);
        for(int i = threadIdx.x; i < P.NRr; i += blockDim.x) L[i] = Rr[i];
        for(int i = threadIdx.x; i < 2 * P.Ntheta; i += blockDim.x) L[P.NRr + i] = trcv[i];
        for(int i = threadIdx.x; i < P.NRz; i += blockDim.x) Lz[i] = Rz[i];
        __syncthreads();
        Rr = L;
        trcv = L + P.NRr;
        Rz = Lz;
    }
    uint32_t n = *O.seg_count;
    if(n > O.seg_cap) n = O.seg_cap;
    Sink3 sink{P.field, 0, &O};
    const Seg3<real> *segs = reinterpret_cast<const Seg3<real> *>(O.segs);
    const uint32_t stride = gridDim.x * blockDim.x;
    for(uint32_t si = blockIdx.x * blockDim.x + threadIdx.x; si < n; si += stride) {
        const Seg3D<real> d = segs[si].d;
        if(d.nitems <= 0) continue;
        const int nr = (d.ir1 - d.ir0) * d.dir + 1;
        for(int j = 0; j < nr; ++j) {
            const int ir = d.ir0 + j * d.dir;
            for(int it = 0; it < P.Ntheta; ++it) {
                real xr, yr, zmin, zmax, wlo, whi;
                if(!row3(d, ir, it, P, Rr, trcv, xr, yr, zmin, zmax)) continue;
                if(!band3_circle(d, xr, yr, zmin, zmax, P, wlo, whi)) continue;
                int i0 = 0, i1 = P.NRz - 1;
                if(P.rz_sorted && wlo == wlo && whi == whi) depth_window(Rz, P.NRz, P.rz_rdz, wlo, whi, i0, i1);
                for(int iz = i0; iz <= i1; ++iz)
                    item3<real, Sink3, FAST, RT>(d, ir, it, iz, xr, yr, zmin, zmax, P, Rz, sink);
            }
        }
    }
    unsigned long long nc = sink.nc;
    for(int off = WAVE / 2; off > 0; off >>= 1) nc += __shfl_xor(nc, off);
    if((threadIdx.x & (WAVE - 1)) == 0 && nc) atomicAdd(O.counters + 1, nc);
}

// Ray run type in 3D: MainRayMode (src/trace.hpp:482-520), one ray per lane,
// every point stored. With a hit list (eigenray re-trace,
// src/mode/eigen.cpp:24-48) ray k starts from hit k and stops after its step.
template<typename real>
__global__ void __launch_bounds__(64) k_ray3d(Prob3D<real> P, bhc_ray_pt3d *pts, int32_t *nsteps, int M, int njobs,
                                            uint32_t *errw, const bhc_eigen_hit *hits)
{
    const int job = blockIdx.x * blockDim.x + threadIdx.x;
    if(job >= njobs) return;
    Err e = {0, 0, 0, 0};
    int ia, ib, isx, isy, isz, lim = -1;
    if(hits) {
        const bhc_eigen_hit h = hits[job];
        ia = h.ialpha;
        ib = h.ibeta;
        isx = h.isx;
        isy = h.isy;
        isz = h.isz;
        lim = h.is;
    } else {
        JobIndices3(job, P, ia, ib, isy, isx, isz);
    }
    bhc_ray_pt3d *out = pts + (size_t)job * M;
    Ray3<real> p0, p1, p2;
    Trace3<real> S;
    Infl3<real> I;
    int N = -1;
    if(!RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e)) {
        store_pt3(out[0], p0);
        N = 1;
    } else {
        store_pt3(out[0], p0);
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdate3(p0, p1, p2, S, P, e);
            store_pt3(out[is + 1], p1);
            if(two) store_pt3(out[is + 2], p2);
            if(lim >= 0 && is >= lim) {
                N = is + 2;
                break;
            }
            is += two ? 2 : 1;
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e, M, N)) break;
        }
        if(N < 0) N = is + 1;
    }
    nsteps[job] = N;
    if(e.err) atomicOr(errw + 0, e.err);
    if(e.warn) atomicOr(errw + 1, e.warn);
    if(e.nerr) atomicAdd(errw + 2, e.nerr);
    if(e.nwarn) atomicAdd(errw + 3, e.nwarn);
}

template<typename real>
int ray3d(bhc_ctx *c, const bhc_problem3d *p, bhc_ray_out3d *out, const bhc_eigen_hit *hits = nullptr,
          int nhits = 0)
{
    std::vector<real> blob;
    Prob3D<real> P;
    std::string perr;
    if(pack3<real>(p, blob, P, perr)) return fail("%s", perr.c_str());
    long long nj64 = hits ? (long long)nhits
                          : (long long)(p->iSingle_alpha >= 1 ? 1 : p->Nalpha) * (p->iSingle_beta >= 1 ? 1 : p->Nbeta)
                                * p->NSx * p->NSy * p->NSz;
    const int M = out->max_points_per_ray;
    if(M < 4) return fail("bhc_ray_run3d: max_points_per_ray must be at least 4");
    if(nj64 * (long long)M > (1ll << 31)) return fail("bhc_ray_run3d: too many ray points for one call");
    const int njobs = (int)nj64;
    hipStream_t s = c->stream;
    real *d_blob = nullptr;
    float *d_Rz = nullptr, *d_Sz = nullptr;
    bhc_ray_pt3d *d_pts = nullptr;
    int32_t *d_ns = nullptr;
    uint32_t *d_errw = nullptr;
    bhc_eigen_hit *d_hits = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        void *ptrs[] = {d_blob, d_Rz, d_Sz, d_pts, d_ns, d_errw, d_hits};
        for(void *q : ptrs)
            if(q) (void)hipFree(q);
    };
    if(dalloc(&d_blob, blob.size()) || dalloc(&d_Rz, (size_t)p->NRz) || dalloc(&d_Sz, (size_t)p->NSz)
       || dalloc(&d_pts, (size_t)njobs * M) || dalloc(&d_ns, (size_t)njobs) || dalloc(&d_errw, 4)
       || (hits && dalloc(&d_hits, (size_t)njobs))) {
        cleanup();
        return -1;
    }
    if(hits && hipMemcpyAsync(d_hits, hits, (size_t)njobs * sizeof(bhc_eigen_hit), hipMemcpyHostToDevice, s) != hipSuccess) {
        cleanup();
        return fail("bhc_ray_run3d_list: upload failed");
    }
    if(hipMemcpyAsync(d_blob, blob.data(), blob.size() * sizeof(real), hipMemcpyHostToDevice, s) != hipSuccess
       || hipMemcpyAsync(d_Rz, p->Rz, p->NRz * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess
       || hipMemcpyAsync(d_Sz, p->Sz, p->NSz * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess
       || hipMemsetAsync(d_errw, 0, 4 * sizeof(uint32_t), s) != hipSuccess) {
        cleanup();
        return fail("bhc_ray_run3d: upload failed");
    }
    P.base = d_blob;
    P.s.base = d_blob;
    P.Rz = d_Rz;
    P.Sz = d_Sz;
    P.njobs = njobs;
    hipLaunchKernelGGL((k_ray3d<real>), dim3((njobs + WAVE - 1) / WAVE), dim3(WAVE), 0, s, P, d_pts, d_ns, M, njobs,
                       d_errw, d_hits);
    if(hipGetLastError() != hipSuccess) rc = fail("bhc_ray_run3d: launch failed");
    uint32_t ew[4] = {0, 0, 0, 0};
    if(!rc && (hipMemcpyAsync(out->points, d_pts, (size_t)njobs * M * sizeof(bhc_ray_pt3d), hipMemcpyDeviceToHost, s)
                   != hipSuccess
               || hipMemcpyAsync(out->nsteps, d_ns, (size_t)njobs * sizeof(int32_t), hipMemcpyDeviceToHost, s)
                   != hipSuccess
               || hipMemcpyAsync(ew, d_errw, sizeof(ew), hipMemcpyDeviceToHost, s) != hipSuccess
               || hipStreamSynchronize(s) != hipSuccess))
        rc = fail("bhc_ray_run3d: download failed");
    out->err = ew[0];
    out->warn = ew[1];
    out->err_count = ew[2];
    out->warn_count = ew[3];
    cleanup();
    return rc;
}

size_t seg3_size(int precision) { return precision == 64 ? sizeof(Seg3<double>) : sizeof(Seg3<float>); }

template<typename real> int run3d(bhc_ctx *c, const bhc_problem3d *p, bhc_field_out *out)
{
    std::vector<real> blob;
    Prob3D<real> P;
    std::string perr;
    if(pack3<real>(p, blob, P, perr)) return fail("%s", perr.c_str());
    long long nj64 = 1;
    nj64 *= (p->iSingle_alpha >= 1 ? 1 : p->Nalpha);
    nj64 *= (p->iSingle_beta >= 1 ? 1 : p->Nbeta);
    nj64 *= (long long)p->NSx * p->NSy * p->NSz;
    if(nj64 >= (1ll << 30)) return fail("3D problem: too many rays");
    const int njobs = (int)nj64;
    const size_t nrcv = (size_t)p->NSz * p->NSx * p->NSy * p->Ntheta * p->NRz * p->NRr;
    hipStream_t s = c->stream;
    real *d_blob = nullptr;
    float *d_Rz = nullptr, *d_Sz = nullptr, *d_field = nullptr;
    int32_t *d_stats = nullptr, *d_work = nullptr;
    double *d_xend = nullptr;
    unsigned long long *d_cnt = nullptr;
    uint32_t *d_errw = nullptr;
    ArrCand *d_cand = nullptr;
    int32_t *d_ncand = nullptr, *d_narr = nullptr, *d_neigen = nullptr;
    bhc_arrival *d_arr = nullptr;
    bhc_eigen_hit *d_hits = nullptr;
    const char r0 = p->RunType[0];
    const int rt = (r0 == 'A' || r0 == 'a' || r0 == 'V' || r0 == 'v') ? RT_ARR : (r0 == 'E') ? RT_EIG : RT_TL;
    const bool hits_too = (r0 == 'E' || r0 == 'V' || r0 == 'v');
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        void *ptrs[] = {d_blob, d_Rz, d_Sz, d_field, d_stats, d_work, d_xend, d_cnt, d_errw,
                        d_cand, d_ncand, d_narr, d_neigen, d_arr, d_hits};
        for(void *q : ptrs)
            if(q) (void)hipFree(q);
    };
    int rc = 0;
#define TRY(x) \
    do { \
        if((x) != 0) { \
            cleanup(); \
            return -1; \
        } \
    } while(0)
#define HTRY(x) \
    do { \
        hipError_t _e = (x); \
        if(_e != hipSuccess) { \
            fail("%s failed: %s", #x, hipGetErrorString(_e)); \
            cleanup(); \
            return -1; \
        } \
    } while(0)
    TRY(dalloc(&d_blob, blob.size()));
    TRY(dalloc(&d_Rz, (size_t)p->NRz));
    TRY(dalloc(&d_Sz, (size_t)p->NSz));
    TRY(dalloc(&d_field, 2 * nrcv));
    TRY(dalloc(&d_work, 1));
    TRY(dalloc(&d_cnt, 2));
    TRY(dalloc(&d_errw, 4));
    if(out->ray_nsteps) {
        TRY(dalloc(&d_stats, 3 * (size_t)njobs));
        TRY(dalloc(&d_xend, 3 * (size_t)njobs));
    }
    HTRY(hipMemcpyAsync(d_blob, blob.data(), blob.size() * sizeof(real), hipMemcpyHostToDevice, s));
    HTRY(hipMemcpyAsync(d_Rz, p->Rz, p->NRz * sizeof(float), hipMemcpyHostToDevice, s));
    HTRY(hipMemcpyAsync(d_Sz, p->Sz, p->NSz * sizeof(float), hipMemcpyHostToDevice, s));
    HTRY(hipMemsetAsync(d_field, 0, 2 * nrcv * sizeof(float), s));
    HTRY(hipMemsetAsync(d_cnt, 0, 2 * sizeof(unsigned long long), s));
    HTRY(hipMemsetAsync(d_errw, 0, 4 * sizeof(uint32_t), s));
    int cand_cap = 0;
    if(rt == RT_ARR) {
        // candidate slots per receiver: 16 GiB worth, raised to 8 x MaxNArr when
        // that still fits 64 GiB; a receiver with more candidates than slots
        // keeps the first ones and raises the out-of-memory warning
        size_t per = nrcv * sizeof(ArrCand);
        size_t cap = std::max<size_t>(1, std::min<size_t>(((size_t)16 << 30) / per, 65536));
        size_t want = 8 * (size_t)out->max_narr;
        if(cap < want && want * per <= ((size_t)64 << 30)) cap = want;
        cand_cap = (int)cap;
        TRY(dalloc(&d_cand, nrcv * cap));
        TRY(dalloc(&d_ncand, nrcv));
        TRY(dalloc(&d_arr, nrcv * (size_t)out->max_narr));
        TRY(dalloc(&d_narr, nrcv));
        HTRY(hipMemsetAsync(d_ncand, 0, nrcv * sizeof(int32_t), s));
    }
    if(hits_too) {
        TRY(dalloc(&d_hits, (size_t)out->eigen_memsize));
        TRY(dalloc(&d_neigen, 1));
        HTRY(hipMemsetAsync(d_neigen, 0, sizeof(int32_t), s));
    }
    P.base = d_blob;
    P.s.base = d_blob;
    P.Rz = d_Rz;
    P.Sz = d_Sz;
    P.field = d_field;
    P.njobs = njobs;
    P.stage_tables = ((size_t)p->NRr + 2 * (size_t)p->Ntheta) * sizeof(real) + (size_t)p->NRz * sizeof(float) <= 56 * 1024;
    TRY(ensure_segs(c));
    DevOut3 O;
    memset(&O, 0, sizeof(O));
    O.stats = d_stats;
    O.xend = d_xend;
    O.counters = d_cnt;
    O.errw = d_errw;
    O.work = d_work;
    O.segs = c->segs;
    O.seg_count = c->d_seg_count;
    O.overflow = c->d_overflow;
    const bool rcen = P.isRayCen != 0;
    O.seg_cap = (uint32_t)std::min<size_t>(c->seg_bytes / (rcen ? sizeof(Seg3RW<real>) : sizeof(Seg3<real>)),
                                           0xFFFFFFF0u);
    O.cand = d_cand;
    O.ncand = d_ncand;
    O.cand_cap = cand_cap;
    O.hits = d_hits;
    O.neigen = d_neigen;
    O.eigen_memsize = hits_too ? out->eigen_memsize : 0;
    O.prog = prog_of(c);
    TRY(prog_begin(c, njobs));
    // split the rays into passes that fit the segment buffer, from the
    // segments per ray of earlier runs on this context (a pass that still
    // overflows is halved and retraced)
    std::vector<std::pair<int, int>> todo;
    {
        int chunk = njobs;
        if(c->seg_per_ray3 > 0) {
            double fit = (double)O.seg_cap / (c->seg_per_ray3 * 1.1 + (double)SEG_BLOCK / WAVE);
            chunk = (int)std::max(1.0, std::min((double)njobs, fit));
        }
        int np = (njobs + chunk - 1) / chunk;
        for(int k = np - 1; k >= 0; --k) todo.push_back({(int)((long long)k * njobs / np), (int)((long long)(k + 1) * njobs / np)});
    }
    c->passes = 0;
    c->ph_trace_ms = c->ph_scatter_ms = 0.0;
    bool first = true;
    unsigned long long segs_total = 0, steps_done = 0;
    while(!todo.empty()) {
        auto r = todo.back();
        todo.pop_back();
        O.job_begin = r.first;
        O.job_end = r.second;
        HTRY(hipMemsetAsync(c->d_seg_count, 0, sizeof(uint32_t), s));
        HTRY(hipMemsetAsync(c->d_overflow, 0, sizeof(uint32_t), s));
        HTRY(hipMemsetAsync(d_work, 0, sizeof(int32_t), s));
        int nj = r.second - r.first;
        int waves = (nj + WAVE - 1) / WAVE;
        int tblocks = std::max(1, std::min(waves, c->num_cus * 4 * c->bpc));
        // error / warning words as the completed passes left them, restored
        // if this trace pass overflows and its rays are traced again
        uint32_t ew_start[4] = {0u, 0u, 0u, 0u};
        unsigned long long prog_start = 0;
        HTRY(hipMemcpyAsync(ew_start, d_errw, sizeof(ew_start), hipMemcpyDeviceToHost, s));
        HTRY(hipMemcpyAsync(&prog_start, c->d_prog, sizeof(prog_start), hipMemcpyDeviceToHost, s));
        if(first) HTRY(hipEventRecord(c->ev0, s));
        HTRY(hipEventRecord(c->pa, s));
        if(rcen) hipLaunchKernelGGL((k_trace3d<real, true>), dim3(tblocks), dim3(WAVE), 0, s, P, O);
        else hipLaunchKernelGGL((k_trace3d<real, false>), dim3(tblocks), dim3(WAVE), 0, s, P, O);
        HTRY(hipGetLastError());
        HTRY(hipEventRecord(c->pb, s));
        uint32_t of = 0;
        unsigned long long steps_now = 0;
        HTRY(hipMemcpyAsync(&of, c->d_overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HTRY(hipMemcpyAsync(&steps_now, d_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HTRY(hipStreamSynchronize(s));
        TRY(phase_add(c, &c->ph_trace_ms));
        if(of) {
            if(nj <= 1) {
                fail("segment buffer too small for a single 3D ray (raise BHC_SEG_BUFFER_MB)");
                cleanup();
                return -1;
            }
            int mid = r.first + nj / 2;
            todo.push_back({mid, r.second});
            todo.push_back({r.first, mid});
            // drop the abandoned pass's steps and warnings, keep the completed
            // passes' (the retried rays report theirs again)
            HTRY(hipMemcpyAsync(d_cnt, &steps_done, sizeof(unsigned long long), hipMemcpyHostToDevice, s));
            HTRY(hipMemcpyAsync(d_errw, ew_start, sizeof(ew_start), hipMemcpyHostToDevice, s));
            HTRY(hipMemcpyAsync(c->d_prog, &prog_start, sizeof(prog_start), hipMemcpyHostToDevice, s));
            HTRY(hipStreamSynchronize(s));
            continue;
        }
        steps_done = steps_now;
        {
            uint32_t ns = 0;
            HTRY(hipMemcpy(&ns, c->d_seg_count, sizeof(uint32_t), hipMemcpyDeviceToHost));
            segs_total += ns;
        }
        if(first) HTRY(hipEventRecord(c->evm, s));
        HTRY(hipEventRecord(c->pa, s));
        int sblocks = c->num_cus * 8;
        size_t lds = P.stage_tables ? ((size_t)P.NRr + 2 * (size_t)P.Ntheta) * sizeof(real) + (size_t)P.NRz * sizeof(float) : 0;
        if(rcen) {
            if(rt == RT_ARR) hipLaunchKernelGGL((k_scatter3rc<real, RT_ARR>), dim3(sblocks), dim3(256), 0, s, P, O);
            else if(rt == RT_EIG) hipLaunchKernelGGL((k_scatter3rc<real, RT_EIG>), dim3(sblocks), dim3(256), 0, s, P, O);
            else hipLaunchKernelGGL((k_scatter3rc<real, RT_TL>), dim3(sblocks), dim3(256), 0, s, P, O);
        } else if(c->scatter3_map == 2 && P.stage_tables &&
                  P.Ntheta < (1 << 26) && (long long)P.NRz * P.NRr < (1ll << 31) &&
                  lds + 4 * ((size_t)SCAT3_G * (sizeof(Seg3<real>) + sizeof(unsigned long long)) + 4 * WAVE * sizeof(int2))
                      <= (size_t)c->lds_max) {
            if(rt == RT_ARR) hipLaunchKernelGGL((k_scatter3d_flat<real, false, RT_ARR>), dim3(sblocks), dim3(256), lds, s, P, O);
            else if(rt == RT_EIG) hipLaunchKernelGGL((k_scatter3d_flat<real, false, RT_EIG>), dim3(sblocks), dim3(256), lds, s, P, O);
            else if(c->exact_phasor) hipLaunchKernelGGL((k_scatter3d_flat<real, false, RT_TL>), dim3(sblocks), dim3(256), lds, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d_flat<real, true, RT_TL>), dim3(sblocks), dim3(256), lds, s, P, O);
        } else if(rt == RT_ARR) {
            if(c->scatter3_map == 1) hipLaunchKernelGGL((k_scatter3d_lanes<real, false, RT_ARR>), dim3(sblocks), dim3(256), lds, s, P, O);
            else if(P.stage_tables) hipLaunchKernelGGL((k_scatter3d<real, false, RT_ARR, true>), dim3(sblocks), dim3(256), lds, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d<real, false, RT_ARR, false>), dim3(sblocks), dim3(256), 0, s, P, O);
        } else if(rt == RT_EIG) {
            if(c->scatter3_map == 1) hipLaunchKernelGGL((k_scatter3d_lanes<real, false, RT_EIG>), dim3(sblocks), dim3(256), lds, s, P, O);
            else if(P.stage_tables) hipLaunchKernelGGL((k_scatter3d<real, false, RT_EIG, true>), dim3(sblocks), dim3(256), lds, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d<real, false, RT_EIG, false>), dim3(sblocks), dim3(256), 0, s, P, O);
        } else if(c->scatter3_map == 1) {
            if(c->exact_phasor) hipLaunchKernelGGL((k_scatter3d_lanes<real, false>), dim3(sblocks), dim3(256), lds, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d_lanes<real, true>), dim3(sblocks), dim3(256), lds, s, P, O);
        } else if(P.stage_tables) {
            if(c->exact_phasor) hipLaunchKernelGGL((k_scatter3d<real, false, RT_TL, true>), dim3(sblocks), dim3(256), lds, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d<real, true, RT_TL, true>), dim3(sblocks), dim3(256), lds, s, P, O);
        } else {
            if(c->exact_phasor) hipLaunchKernelGGL((k_scatter3d<real, false, RT_TL, false>), dim3(sblocks), dim3(256), 0, s, P, O);
            else hipLaunchKernelGGL((k_scatter3d<real, true, RT_TL, false>), dim3(sblocks), dim3(256), 0, s, P, O);
        }
        HTRY(hipGetLastError());
        HTRY(hipEventRecord(c->pb, s));
        TRY(phase_add(c, &c->ph_scatter_ms));
        ++c->passes;
        first = false;
    }
    if(rt == RT_ARR) {
        // AddArr in the single-threaded order (or the first MaxNArr without merging)
        const double om = 2.0 * M_PI * p->freq0;
        const float omf = 2.0f * (float)M_PI * (float)p->freq0;
        hipLaunchKernelGGL(k_arr_merge, dim3((unsigned)((nrcv + 255) / 256)), dim3(256), 0, s, d_cand, d_ncand, cand_cap,
                           d_arr, d_narr, out->max_narr, (int)nrcv, out->allow_merging, omf, om,
                           sizeof(real) == 8 ? 1 : 0, d_errw);
        HTRY(hipGetLastError());
    }
    HTRY(hipEventRecord(c->ev1, s));
    c->seg_per_ray3 = std::max(c->seg_per_ray3, (double)segs_total / (double)njobs);
    if(rt == RT_TL && out->field)
        HTRY(hipMemcpyAsync(out->field, d_field, 2 * nrcv * sizeof(float), hipMemcpyDeviceToHost, s));
    if(rt == RT_ARR) {
        HTRY(hipMemcpyAsync(out->arr, d_arr, nrcv * (size_t)out->max_narr * sizeof(bhc_arrival), hipMemcpyDeviceToHost,
                            s));
        HTRY(hipMemcpyAsync(out->narr, d_narr, nrcv * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    int32_t ne = 0;
    if(hits_too) {
        HTRY(hipMemcpyAsync(&ne, d_neigen, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HTRY(hipStreamSynchronize(s));
        *out->neigen = ne;
        int nh = std::min(ne, out->eigen_memsize);
        if(nh > 0)
            HTRY(hipMemcpyAsync(out->hits, d_hits, (size_t)nh * sizeof(bhc_eigen_hit), hipMemcpyDeviceToHost, s));
    }
    uint32_t ew[4];
    unsigned long long cnt[2];
    HTRY(hipMemcpyAsync(ew, d_errw, sizeof(ew), hipMemcpyDeviceToHost, s));
    HTRY(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, s));
    std::vector<int32_t> st;
    if(d_stats) {
        st.resize(3 * (size_t)njobs);
        HTRY(hipMemcpyAsync(st.data(), d_stats, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        if(out->ray_xend)
            HTRY(hipMemcpyAsync(out->ray_xend, d_xend, 3 * (size_t)njobs * sizeof(double), hipMemcpyDeviceToHost,
                                s));
    }
    HTRY(hipStreamSynchronize(s));
    out->err = ew[0];
    out->warn = ew[1];
    out->err_count = ew[2];
    out->warn_count = ew[3];
    out->total_steps = (int64_t)cnt[0];
    out->total_contrib = (int64_t)cnt[1];
    if(d_stats)
        for(int j = 0; j < njobs; ++j) {
            out->ray_nsteps[j] = st[3 * j];
            if(out->ray_ntop) out->ray_ntop[j] = st[3 * j + 1];
            if(out->ray_nbot) out->ray_nbot[j] = st[3 * j + 2];
        }
    cleanup();
#undef TRY
#undef HTRY
    return rc;
}

} // namespace

extern "C" int bhc_field_run3d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out)
{
    if(check_idle(ctx, "bhc_field_run3d")) return -1;
    if(!prob || !out) return fail("bhc_field_run3d: null argument");
    if(prob->precision != 64 && prob->precision != 32) return fail("bhc_field_run3d: precision must be 32 or 64");
    if(prob->RunType[5] == '2') return fail("bhc_field_run3d: an Nx2D problem (RunType[5] '2') runs with bhc_field_runnx2d");
    char r0 = prob->RunType[0], r1 = prob->RunType[1];
    if(r0 != 'C' && r0 != 'S' && r0 != 'I' && r0 != 'A' && r0 != 'a' && r0 != 'E' && r0 != 'V' && r0 != 'v')
        return fail("bhc_field_run3d: run type '%c' is not a TL, arrivals or eigenray run type", r0);
    if((r0 == 'A' || r0 == 'a' || r0 == 'V' || r0 == 'v') && (!out->arr || !out->narr || out->max_narr < 1))
        return fail("bhc_field_run3d: arrivals run needs arr, narr and max_narr >= 1");
    if((r0 == 'E' || r0 == 'V' || r0 == 'v') && (!out->hits || !out->neigen || out->eigen_memsize < 1))
        return fail("bhc_field_run3d: eigenray run needs hits, neigen and eigen_memsize >= 1");
    if(r1 != 'G' && r1 != '^' && r1 != ' ' && r1 != 'B' && r1 != 'g' && r1 != 'b')
        return fail("bhc_field_run3d: beam type '%c' is not a geometric beam", r1);
    if((r1 == 'g' || r1 == 'b') && prob->NRr > 1 && !(prob->Delta_r > 0))
        return fail("bhc_field_run3d: ray-centred beams need equally spaced receiver ranges");
    if(prob->NSx < 1 || prob->NSy < 1 || prob->NSz < 1 || prob->NRz < 1 || prob->NRr < 1 || prob->Ntheta < 1
       || prob->Nalpha < 1 || prob->Nbeta < 1)
        return fail("bhc_field_run3d: empty source / receiver / angle set");
    if(prob->NSBPPts < 2) return fail("bhc_field_run3d: source beam pattern needs 2 points");
    if(!(prob->deltas > 0)) return fail("bhc_field_run3d: deltas must be positive");
    if(hipSetDevice(ctx->device) != hipSuccess) return fail("bhc_field_run3d: hipSetDevice failed");
    const int rc = prob->precision == 64 ? run3d<double>(ctx, prob, out) : run3d<float>(ctx, prob, out);
    if(!rc) ctx->prog_final = 1;
    return rc;
}

extern "C" int bhc_ray_run3d(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_ray_out3d *out)
{
    if(check_idle(ctx, "bhc_ray_run3d")) return -1;
    if(!prob || !out || !out->points || !out->nsteps) return fail("bhc_ray_run3d: null argument");
    if(prob->RunType[5] == '2') return fail("bhc_ray_run3d: an Nx2D problem (RunType[5] '2') runs with bhc_ray_runnx2d");
    if(prob->NSx < 1 || prob->NSy < 1 || prob->NSz < 1 || prob->Nalpha < 1 || prob->Nbeta < 1)
        return fail("bhc_ray_run3d: empty source / angle set");
    if(!(prob->deltas > 0)) return fail("bhc_ray_run3d: deltas must be positive");
    if(hipSetDevice(ctx->device) != hipSuccess) return fail("bhc_ray_run3d: hipSetDevice failed");
    const long long n = (long long)(prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta)
        * prob->NSx * prob->NSy * prob->NSz;
    if(prog_begin(ctx, n)) return -1;
    const int rc = prob->precision == 64 ? ray3d<double>(ctx, prob, out) : ray3d<float>(ctx, prob, out);
    if(!rc) ctx->prog_final = 1;
    return rc;
}

extern "C" int bhc_ray_run3d_list(bhc_ctx *ctx, const bhc_problem3d *prob, const bhc_eigen_hit *hits, int32_t nhits,
                                  bhc_ray_out3d *out)
{
    if(check_idle(ctx, "bhc_ray_run3d_list")) return -1;
    if(!prob || !hits || !out || !out->points || !out->nsteps) return fail("bhc_ray_run3d_list: null argument");
    if(nhits < 1) return fail("bhc_ray_run3d_list: no hits");
    if(prob->RunType[5] == '2') return fail("bhc_ray_run3d_list: Nx2D problems are not supported");
    for(int k = 0; k < nhits; ++k) {
        const bhc_eigen_hit &h = hits[k];
        if(h.ialpha < 0 || h.ialpha >= prob->Nalpha || h.ibeta < 0 || h.ibeta >= prob->Nbeta || h.isx < 0
           || h.isx >= prob->NSx || h.isy < 0 || h.isy >= prob->NSy || h.isz < 0 || h.isz >= prob->NSz || h.is < 0)
            return fail("bhc_ray_run3d_list: hit %d names a ray outside the problem", k);
    }
    if(!(prob->deltas > 0)) return fail("bhc_ray_run3d_list: deltas must be positive");
    if(hipSetDevice(ctx->device) != hipSuccess) return fail("bhc_ray_run3d_list: hipSetDevice failed");
    if(prog_begin(ctx, nhits)) return -1;
    const int rc = prob->precision == 64 ? ray3d<double>(ctx, prob, out, hits, nhits)
                                         : ray3d<float>(ctx, prob, out, hits, nhits);
    if(!rc) ctx->prog_final = 1;
    return rc;
}

// ---------------------------------------------------------------------------
// Non-blocking runs (the reference's extsetup_blocking(params, false) mode,
// src/mode/fieldimpl.cpp.in:46-52,66-71): the run is performed by a worker
// thread of the context, which fires the completion callback once the outputs
// have been written; bhc_progress reports its finished rays meanwhile.
// ---------------------------------------------------------------------------
namespace {
template<typename F> int start_async(bhc_ctx *ctx, const char *who, long long total, F run, bhc_completed_fn cb,
                                     void *user)
{
    if(!ctx) return fail("%s: null context", who);
    if(tl_async_worker) return fail("%s: a run cannot be started from a completion callback", who);
    int idle = 0;
    if(!ctx->busy.compare_exchange_strong(idle, 1)) return fail("%s: a non-blocking run is in flight", who);
    if(ctx->worker.joinable()) ctx->worker.join();   // the previous run has finished
    ctx->prog_final = 0;
    ctx->prog_seen = 0;
    ctx->prog_total = total;
    __atomic_store_n(ctx->h_prog, 0ull, __ATOMIC_RELAXED);
    ctx->async_rc = 0;
    ctx->async_msg.clear();
    try {
        ctx->worker = std::thread([ctx, run, cb, user]() {
            tl_async_worker = true;
            const int rc = run();
            ctx->async_rc = rc;
            if(rc) ctx->async_msg = g_last_error;
            ctx->busy = 0;
            if(cb) cb(user);
        });
    } catch(...) {
        ctx->busy = 0;
        return fail("%s: could not start the worker thread", who);
    }
    return 0;
}
} // namespace

extern "C" int bhc_field_run2d_async(bhc_ctx *ctx, const bhc_problem2d *prob, bhc_field_out *out,
                                     bhc_completed_fn completed, void *user)
{
    if(!prob || !out) return fail("bhc_field_run2d_async: null argument");
    if(check_problem(prob)) return -1;
    const long long total = (long long)(prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * prob->NSz;
    return start_async(ctx, "bhc_field_run2d_async", total, [ctx, prob, out]() { return bhc_field_run2d(ctx, prob, out); },
                       completed, user);
}

extern "C" int bhc_field_run3d_async(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out,
                                     bhc_completed_fn completed, void *user)
{
    if(!prob || !out) return fail("bhc_field_run3d_async: null argument");
    const long long total = (long long)(prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha)
        * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta) * prob->NSx * prob->NSy * prob->NSz;
    return start_async(ctx, "bhc_field_run3d_async", total, [ctx, prob, out]() { return bhc_field_run3d(ctx, prob, out); },
                       completed, user);
}

extern "C" int bhc_field_runnx2d_async(bhc_ctx *ctx, const bhc_problem3d *prob, bhc_field_out *out,
                                       bhc_completed_fn completed, void *user)
{
    if(!prob || !out) return fail("bhc_field_runnx2d_async: null argument");
    if(prob->RunType[5] != '2') return fail("bhc_field_runnx2d_async: not an Nx2D problem (RunType[5] must be '2')");
    const long long total = (long long)(prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha)
        * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta) * prob->NSx * prob->NSy * prob->NSz;
    return start_async(ctx, "bhc_field_runnx2d_async", total,
                       [ctx, prob, out]() { return bhc_field_runnx2d(ctx, prob, out); }, completed, user);
}

extern "C" int bhc_ctx_busy(bhc_ctx *ctx) { return ctx && ctx->busy.load() ? 1 : 0; }

extern "C" int bhc_ctx_wait(bhc_ctx *ctx)
{
    if(!ctx) return fail("bhc_ctx_wait: null context");
    if(tl_async_worker) return fail("bhc_ctx_wait: cannot wait from a completion callback");
    if(ctx->worker.joinable()) ctx->worker.join();
    if(ctx->async_rc) {
        g_last_error = ctx->async_msg;
        return ctx->async_rc;
    }
    return 0;
}
