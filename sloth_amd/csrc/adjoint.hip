// k4: adjoint of the RK4 trace, for d(loss on the final ray states)/d(n_e).
//
// Discrete adjoint: the gradient of exactly the fixed-step RK4 map the
// forward kernel integrates (same schedule, same straight-line moves), so a
// central finite difference of the forward loss converges to it.
//
//  * slt_forward_ckpt runs the forward integration and stores each ray's state
//    every K steps (a checkpoint), plus the final state and tau.
//  * slt_adjoint walks the checkpoint intervals from last to first: it replays
//    the interval from its checkpoint onto a small per-ray tape (position,
//    velocity and step length of every move), then runs the transposed RK4
//    step over the tape in reverse.  Each stage re-evaluates the cached cell's
//    trilinear polynomial; its Jacobian gives the state adjoint, and the
//    stage's share of the grid adjoint is accumulated per lane as the 8
//    polynomial moments of the current cell (in registers within a step, in
//    the lane's LDS row between steps) and flushed to global memory (atomics,
//    combined across the lanes of a cell) only when the cell changes.
//  * With ckpt_every = 0 there are no checkpoints: slt_forward_ckpt records
//    every move of the whole trajectory onto the caller's tape (n_rays x
//    sched_steps records at most, memory traded for the replay), and
//    slt_adjoint reads that tape in reverse without replaying anything.
//  * slt_adjoint_unpack turns the voxel adjoint (d/dx, d/dy, d/dz of coeff ne,
//    kappa) into dL/dn_e: the transposed numpy.gradient stencil times coeff,
//    plus d kappa / d n_e times the kappa adjoint.
#include <math.h>

#include <mutex>

#include "integrate.h"

namespace slt {

// ---------------------------------------------------------------------------
// Checkpoints: per checkpoint, SoA over rays:
//   rb[3] f64 | tau f64 | off[3] f32 | dv[3] f32 | fold u32      (60 B per ray)
// ---------------------------------------------------------------------------
constexpr int64_t kCkptBytesPerRay = 4 * 8 + 6 * 4 + 4;

struct CkptView {
    double* rb[3];
    double* tau;
    float* off[3];
    float* dv[3];
    uint32_t* fold;
};

__device__ __forceinline__ CkptView ckpt_view(void* base, int64_t n, int c) {
    CkptView v;
    char* p = static_cast<char*>(base) + (int64_t)c * kCkptBytesPerRay * n;
    double* d = reinterpret_cast<double*>(p);
    for (int a = 0; a < 3; ++a) v.rb[a] = d + a * n;
    v.tau = d + 3 * n;
    float* f = reinterpret_cast<float*>(d + 4 * n);
    for (int a = 0; a < 3; ++a) {
        v.off[a] = f + a * n;
        v.dv[a] = f + (3 + a) * n;
    }
    v.fold = reinterpret_cast<uint32_t*>(f + 6 * n);
    return v;
}

// Tape: per slot (one move), one 32-byte record per ray: (rh[3], hT) and
// (vh[3], unused), hT = h > 0 for an RK4 step, -T for a straight move of
// duration T (whose record holds only hT).  A slot is two planes of 16-byte
// halves, every ray's first half and then every ray's second half, so that each
// 16-byte store or load of a wavefront covers 1 KiB without gaps (records of 32
// contiguous bytes left every instruction half of each line: 11.7 vs 10.3 ms
// for the trajectory-recording forward on 256^3 / 2e6 rays).
constexpr int64_t kTapeRecBytes = 32;

struct GlobalTape {
    float* base;
    int64_t n;
    int64_t t;
    int count;
    __device__ __forceinline__ slt_f4* rec(int j) const {
        return reinterpret_cast<slt_f4*>(base) + ((int64_t)j * 2 * n + t);
    }
    __device__ __forceinline__ void rk(const float* rh, const float* vh, float h) {
        slt_f4* p = rec(count++);
        p[0] = slt_f4{rh[0], rh[1], rh[2], h};
        p[n] = slt_f4{vh[0], vh[1], vh[2], 0.f};
    }
    __device__ __forceinline__ void straight(float T) {
        slt_f4* p = rec(count++);
        p[0] = slt_f4{0.f, 0.f, 0.f, -T};
    }
};

struct AdjArgs {
    DevGrid G;
    int64_t n;
    const float* __restrict__ s0;
    const int32_t* __restrict__ order;
    int nseg;
    int K;
    int nck;
    int64_t sched;
    const int32_t* __restrict__ seg_nsub;
    const float* __restrict__ seg_h;
    void* ckpt;
    float* __restrict__ final_state;
    double* __restrict__ final_tau;
    const float* __restrict__ seed_state;
    const double* __restrict__ seed_tau;
    float* __restrict__ grid_adj;
    float* __restrict__ tape;
    unsigned long long* __restrict__ total;
    int nblocks;
};

// Five wavefronts per SIMD, as the offline trace kernel: at 96 VGPRs the
// values that spill are touched around the save points and in the rare
// no-stage-in-the-box branch, not in the step loop (7.9 vs 8.3 ms at four).
// TAPE (ckpt_every = 0): no checkpoints; every move of the whole trajectory
// goes onto the tape and each ray's record count to the checkpoint area.
template <bool KAPPA, bool TAPE, bool UNI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
forward_ckpt_kernel(AdjArgs A) {
    const int lb = xcd_block(blockIdx.x, gridDim.x, A.nblocks);
    if (lb < 0) return;
    const int64_t t = (int64_t)lb * blockDim.x + threadIdx.x;
    uint32_t nsteps = 0;
    if (t < A.n) {
        const int64_t ray = A.order ? (int64_t)A.order[t] : t;
        RayState s;
        state_from_rays(A.G, A.s0, A.n, ray, s);
        SaveSink sink{A.final_state, A.final_tau, A.n, ray, 0, A.nseg, A.s0 + 3 * A.n, A.n, ray};
        if (TAPE) {
            // plain stores: a load's wait also waits for the stores before it, and
            // non-temporal stores are acknowledged later (15.7 vs 11.5 ms on
            // 256^3 / 2e6 rays)
            GlobalTape tape{A.tape, A.n, t, 0};
            integrate<KAPPA, false, GlobalTape, UNI>(A.G, s, A.seg_nsub, A.seg_h, A.nseg, 0,
                                                     A.sched, sink, tape);
            static_cast<int32_t*>(A.ckpt)[t] = tape.count;
        }
        for (int c = 0; !TAPE && c < A.nck; ++c) {
            const CkptView v = ckpt_view(A.ckpt, A.n, c);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                v.rb[a][t] = s.rb[a];
                v.off[a][t] = s.off[a];
                v.dv[a][t] = s.dv[a];
            }
            v.tau[t] = s.tau;
            v.fold[t] = s.fold;
            const int64_t g0 = (int64_t)c * A.K;
            const int64_t g1 = g0 + A.K < A.sched ? g0 + A.K : A.sched;
            integrate<KAPPA, false, UNI>(A.G, s, A.seg_nsub, A.seg_h, A.nseg, g0, g1, sink);
        }
        nsteps = inbox_steps(s);
    }
    if (A.total) add_steps(A.total, nsteps);
}

// Per-lane grid-adjoint accumulator: the 8 polynomial moments (per channel) of
// the stage adjoints of this round that fell into the lane's cell; the moments
// of earlier rounds in the same cell are in the lane's LDS row.
struct Moments {
    int id;  // cell (voxel index of corner (i, j, k)); -1 = empty
    float4 m0, mx, my, mz, mxy, mxz, myz, mxyz;
};

__device__ __forceinline__ void f4fma(float4& acc, float s, const float4& a) {
    acc.x = fmaf(s, a.x, acc.x);
    acc.y = fmaf(s, a.y, acc.y);
    acc.z = fmaf(s, a.z, acc.z);
    acc.w = fmaf(s, a.w, acc.w);
}

__device__ __forceinline__ float4 f4add(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

#define SLT_LDS __attribute__((address_space(3)))

// Corner adjoints of one lane's moments, read back from its LDS row (8 float4:
// m0, mx, my, mz, mxy, mxz, myz, mxyz) one channel at a time: corner E (axis
// bits x = 4, y = 2, z = 1) = sum over monomial sets S containing E of
// (-1)^|S \ E| m_S, three butterfly steps over the 8 moments; each is added to
// its voxel's adjoint.  Used for the cells of a flush beyond its slots, so the
// moment registers are not read here and nothing of this rare path is kept
// live across the flush's cell loop.
__device__ __forceinline__ void scatter_row(const DevGrid& G, float* __restrict__ adj, int id,
                                            const SLT_LDS float* row) {
#pragma unroll 1
    for (int ch = 0; ch < 4; ++ch) {
        float m[8];
#pragma unroll
        for (int S = 0; S < 8; ++S) m[S] = row[4 * ((0x74516230 >> (4 * S)) & 7) + ch];
#pragma unroll
        for (int b = 1; b < 8; b <<= 1) {
#pragma unroll
            for (int S = 0; S < 8; ++S)
                if (!(S & b)) m[S] -= m[S | b];
        }
#pragma unroll
        for (int S = 0; S < 8; ++S) {
            const int off = ((S >> 2) & 1) * G.s0 + ((S >> 1) & 1) * G.s1 + (S & 1) * G.s2;
            atomicAdd(adj + (int64_t)(id + off) * 4 + ch, m[S]);
        }
    }
}

// Where flushed moments go: the voxel adjoint in global memory and this
// wavefront's flush area in LDS: one row of 32 moments per lane (8 float4: m0,
// mx, my, mz, mxy, mxz, myz, mxyz; rows 16-byte aligned) and a 65th row of
// zeros, then kFlushSlots cell ids, their member counts, their member lists (up
// to 64 row numbers of one byte each per slot) and their corner adjoints (8
// float4 per slot).
constexpr int kFlushSlots = 16;
constexpr int kRowStride = 36;
constexpr int kWaveLds = 65 * kRowStride + 2 * kFlushSlots + kFlushSlots * 16 + kFlushSlots * 32;
static_assert(kRowStride % 4 == 0 && kWaveLds % 4 == 0, "rows must stay 16-byte aligned");

struct AdjOut {
    float* __restrict__ adj;
    SLT_LDS float* lds;
    __device__ __forceinline__ SLT_LDS int* ids() const {
        return reinterpret_cast<SLT_LDS int*>(lds + 65 * kRowStride);
    }
    __device__ __forceinline__ SLT_LDS int* counts() const { return ids() + kFlushSlots; }
    __device__ __forceinline__ SLT_LDS uint32_t* lists() const {
        return reinterpret_cast<SLT_LDS uint32_t*>(counts() + kFlushSlots);
    }
    __device__ __forceinline__ SLT_LDS slt_f4* corners() const {
        return reinterpret_cast<SLT_LDS slt_f4*>(lists() + kFlushSlots * 16);
    }
};

// Adds a lane's moment sums to its LDS row.
__device__ __forceinline__ void add_to_row(const AdjOut& out, const Moments& M) {
    SLT_LDS slt_f4* row = reinterpret_cast<SLT_LDS slt_f4*>(out.lds + __lane_id() * kRowStride);
    const float4* src[8] = {&M.m0, &M.mx, &M.my, &M.mz, &M.mxy, &M.mxz, &M.myz, &M.mxyz};
#pragma unroll
    for (int q = 0; q < 8; ++q) row[q] += slt_f4{src[q]->x, src[q]->y, src[q]->z, src[q]->w};
}

// Orders this wavefront's LDS accesses across its lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Collective flush (whole wavefront, uniform control flow): the lanes with
// `need` set give their moments to their cell.  A sorted wavefront holds a few
// cells at a time, and at a flush nearly all its lanes leave theirs together.
// Every lane adds its round's sums to its LDS row (which holds its moments of
// the current cell), so no moment stays in registers across the flush; a
// scalar loop over the flushing lanes' cells gives each cell a slot and lists
// its member lanes there; then each
// lane sums the member rows of one (slot, monomial) entry, all four channels per
// 16-byte read, four member reads in flight (a fixed order per wavefront), and
// the eight lanes of a slot turn their monomial sums into the eight corner
// adjoints with three butterfly steps.  The corners go through LDS so that each
// atomic instruction adds the four channels of 16 corners: a flush of c cells
// issues ceil(32 c / 64) full-width atomic instructions whose lanes share voxels.
// No LDS atomics: float atomics on LDS addresses shared by many lanes serialise.
// Lanes of cells beyond the slots scatter their own.  The flushing lanes' rows
// are then cleared for their next cells.
template <bool SUMS = true>
__device__ __forceinline__ void flush_wave(const DevGrid& G, const AdjOut& out, Moments& M,
                                           bool need) {
    const int lane = __lane_id();
    // every lane with a cell moves its sums to its row, so no moment is held in
    // registers across the flush
    if (SUMS) {
        if (M.id >= 0) add_to_row(out, M);
        M.m0 = M.mx = M.my = M.mz = M.mxy = M.mxz = M.myz = M.mxyz =
            make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wave_lds_sync();
    // member lists start out naming row 64, which is all zeros, so a sum that reads a
    // whole word of four entries past a list's end adds nothing
    reinterpret_cast<SLT_LDS slt_f4*>(out.lists())[lane] =
        __builtin_bit_cast(slt_f4, make_uint4(0x40404040u, 0x40404040u, 0x40404040u, 0x40404040u));
    wave_lds_sync();
    // per cell (scalar loop): its member lanes' ids in a list, their count, the cell id
    uint64_t pending = __ballot(need);
    int nslots = 0;
    while (pending && nslots < kFlushSlots) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int key = __builtin_amdgcn_readlane(M.id, leader);
        const bool mine = need && M.id == key;
        const uint64_t mem = __ballot(mine);
        pending &= ~mem;
        if (mine) {
            const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(mem >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mem, 0));
            reinterpret_cast<SLT_LDS uint8_t*>(out.lists() + nslots * 16)[pos] = (uint8_t)lane;
        }
        if (lane == 0) {
            out.ids()[nslots] = key;
            out.counts()[nslots] = __popcll(mem);
        }
        ++nslots;
    }
    // lanes of cells beyond the slots scatter their own moments from their rows
    if (pending) {
        if ((pending >> lane) & 1) scatter_row(G, out.adj, M.id, out.lds + lane * kRowStride);
    }
    wave_lds_sync();
    // per (slot, monomial S) entry (lane-parallel), the eight monomials of a slot in
    // eight adjacent lanes: the sum over the slot's member rows of moment q(S), all four
    // channels in one 16-byte read per member, four member reads in flight; then the
    // corner adjoints by the superset Moebius transform across those lanes (corner E =
    // sum over S containing E of (-1)^|S \ E| m_S), one butterfly per axis bit; each lane
    // adds its corner's four channels to the voxel adjoint.
    const SLT_LDS slt_f4* rows = reinterpret_cast<const SLT_LDS slt_f4*>(out.lds);
    for (int e = lane; e < nslots * 8; e += 64) {
        const int sl = e >> 3, S = e & 7;
        const int q = (0x74516230 >> (4 * S)) & 7;  // the moment of monomial set S
        const int cnt = out.counts()[sl];
        const SLT_LDS uint32_t* lst = out.lists() + sl * 16;
        slt_f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < cnt; i += 4) {
            const uint32_t w = lst[i >> 2];
            slt_f4 v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = rows[((w >> (8 * r)) & 127u) * (kRowStride / 4) + q];
            acc += (v[0] + v[1]) + (v[2] + v[3]);
        }
#pragma unroll
        for (int bit = 1; bit < 8; bit <<= 1) {
            const slt_f4 partner = {__shfl_xor(acc.x, bit), __shfl_xor(acc.y, bit),
                                    __shfl_xor(acc.z, bit), __shfl_xor(acc.w, bit)};
            acc = (S & bit) ? acc : acc - partner;
        }
        out.corners()[e] = acc;
    }
    wave_lds_sync();
    // each lane adds one channel of one corner; one instruction covers the 4 channels of
    // 16 corners, so the atomics on a voxel's channels travel together
    for (int e = lane; e < nslots * 32; e += 64) {
        const int sl = e >> 5, ch = (e >> 3) & 3, S = e & 7;
        const int off = ((S >> 2) & 1) * G.s0 + ((S >> 1) & 1) * G.s1 + (S & 1) * G.s2;
        atomicAdd(out.adj + (int64_t)(out.ids()[sl] + off) * 4 + ch,
                  reinterpret_cast<const SLT_LDS float*>(out.corners())[(sl * 8 + S) * 4 + ch]);
    }
    // the lanes that gave their moments start their next cell from zero
    if (need) {
        SLT_LDS slt_f4* row = reinterpret_cast<SLT_LDS slt_f4*>(out.lds + lane * kRowStride);
#pragma unroll
        for (int q = 0; q < 8; ++q) row[q] = slt_f4{0.f, 0.f, 0.f, 0.f};
    }
    wave_lds_sync();
}

// The reverse steps' cell cache: polynomial, voxel index of corner (i, j, k)
// (-1 = none), that corner in grid units (NaN = none) and whether the cell
// touches a face of the grid.
struct AdjCell : CellPoly {
    int id;
    float c0, c1, c2;
    bool face;
    float w0, w1, w2;  // non-uniform grids: 1 / (cell width in grid units) per axis
};

// Locates the grid-unit point u of a lane with a step (act): returns whether it
// lies in the closed box, and its local coordinates t in the cell the cache then
// holds (floor(u), clamped to the last cell on each axis, as the forward finds
// it).  A point in the cached cell, when that cell is interior, is inside and
// needs nothing more (t = u - corner is exact there).  When any lane of the
// wavefront has another point, every lane takes the full test together, and if
// any lane needs a new cell every lane reloads its cell (a lane that keeps its
// cell reloads it; one with no cell loads voxel 0 and keeps id -1), so the old
// polynomial is dead where the new one is written.
template <bool UNI>
__device__ __forceinline__ bool locate(const DevGrid& G, bool act, float u0, float u1, float u2,
                                       AdjCell& cc, float& tx, float& ty, float& tz) {
    tx = UNI ? u0 - cc.c0 : (u0 - cc.c0) * cc.w0;
    ty = UNI ? u1 - cc.c1 : (u1 - cc.c1) * cc.w1;
    tz = UNI ? u2 - cc.c2 : (u2 - cc.c2) * cc.w2;
    const bool fast = (unit_interval(tx) & unit_interval(ty) & unit_interval(tz)) && !cc.face;
    bool in = act;
    if (__ballot(act && !fast)) {
        // the full test runs only in this branch: an empty asm statement hands the point
        // over inside it, so the compiler does not compute the box test and the cell index
        // ahead of the branch for every call
        asm volatile("" : "+v"(u0), "+v"(u1), "+v"(u2));
        const bool inb = in_box(G, u0, u1, u2);
        in = act && (fast || inb);
        const bool moving = act && !fast && inb;
        // the cell as the forward finds it: floor(u) on a uniform grid, the knot search
        // otherwise (only for the lanes that move: the others' points may be outside)
        const int i0 = UNI ? min((int)u0, G.n0 - 2) : (moving ? knot_cell(G.kn0, G.n0, u0) : 0);
        const int i1 = UNI ? min((int)u1, G.n1 - 2) : (moving ? knot_cell(G.kn1, G.n1, u1) : 0);
        const int i2 = UNI ? min((int)u2, G.n2 - 2) : (moving ? knot_cell(G.kn2, G.n2, u2) : 0);
        const int want =
            moving ? (int)(__umul24(i0, G.s0) + __umul24(i1, G.s1) + __umul24(i2, G.s2)) : cc.id;
        if (__ballot(want != cc.id)) {
            load_cell(G, want < 0 ? 0 : want, cc);
            cc.id = want;
        }
        if (moving) {
            if (UNI) {
                cc.c0 = (float)i0;
                cc.c1 = (float)i1;
                cc.c2 = (float)i2;
            } else {
                cc.c0 = G.kn0[i0];
                cc.c1 = G.kn1[i1];
                cc.c2 = G.kn2[i2];
                cc.w0 = 1.0f / (G.kn0[i0 + 1] - cc.c0);
                cc.w1 = 1.0f / (G.kn1[i1 + 1] - cc.c1);
                cc.w2 = 1.0f / (G.kn2[i2 + 1] - cc.c2);
            }
            cc.face = (min(min(i0, i1), i2) == 0) | (i0 == G.n0 - 2) | (i1 == G.n1 - 2) |
                      (i2 == G.n2 - 2);
            tx = UNI ? u0 - cc.c0 : (u0 - cc.c0) * cc.w0;
            ty = UNI ? u1 - cc.c1 : (u1 - cc.c1) * cc.w1;
            tz = UNI ? u2 - cc.c2 : (u2 - cc.c2) * cc.w2;
        }
    }
    return in;
}

// Adds one stage's adjoint Ab at cell `id`, local coordinates t, to the lane's
// moments (contrib = this lane has a contribution).  Called by every lane of
// the wavefront together.
template <bool FIRST>
__device__ __forceinline__ void accumulate(const DevGrid& G, const AdjOut& out, Moments& M,
                                           bool contrib, int id, float tx, float ty, float tz,
                                           const float4& Ab) {
    const bool move = contrib && id != M.id;
    // (the round's first stage holds no sums yet: they went to the rows at the end of
    // the previous round)
    if (__ballot(move && M.id >= 0)) flush_wave<!FIRST>(G, out, M, move && M.id >= 0);
    M.id = move ? id : M.id;
    if (FIRST) {
        // the round's first stage: the sums are zero here (cleared at the end of the
        // previous round and by any flush), so it writes them (Ab = 0 without a
        // contribution)
        const float4 A0 = make_float4(contrib ? Ab.x : 0.f, contrib ? Ab.y : 0.f,
                                      contrib ? Ab.z : 0.f, contrib ? Ab.w : 0.f);
        const float txy = tx * ty, txz = tx * tz, tyz = ty * tz, txyz = txy * tz;
        auto mul = [](float s, const float4& a) {
            return make_float4(s * a.x, s * a.y, s * a.z, s * a.w);
        };
        M.m0 = A0;
        M.mx = mul(tx, A0);
        M.my = mul(ty, A0);
        M.mz = mul(tz, A0);
        M.mxy = mul(txy, A0);
        M.mxz = mul(txz, A0);
        M.myz = mul(tyz, A0);
        M.mxyz = mul(txyz, A0);
    } else if (contrib) {
        const float txy = tx * ty, txz = tx * tz, tyz = ty * tz, txyz = txy * tz;
        M.m0 = f4add(M.m0, Ab);
        f4fma(M.mx, tx, Ab);
        f4fma(M.my, ty, Ab);
        f4fma(M.mz, tz, Ab);
        f4fma(M.mxy, txy, Ab);
        f4fma(M.mxz, txz, Ab);
        f4fma(M.myz, tyz, Ab);
        f4fma(M.mxyz, txyz, Ab);
    }
}

// Adjoint of one RHS evaluation at the grid-unit point u with output adjoint Ab
// (gx, gy, gz, kappa): returns d(loss)/d(u) in rb and accumulates the grid
// adjoint.  Every lane of the wavefront calls it together; `act` lanes contribute.
template <bool FIRST, bool UNI>
__device__ __forceinline__ void eval_adj(const DevGrid& G, bool act, float u0, float u1, float u2,
                                         const float4& Ab, AdjCell& cc, Moments& M,
                                         const AdjOut& out, float rb[3]) {
    rb[0] = rb[1] = rb[2] = 0.f;
    float tx, ty, tz;
    const bool in = locate<UNI>(G, act, u0, u1, u2, cc, tx, ty, tz);
    if (in) {
        // channel-contracted coefficients: Kq = sum_ch Ab_ch * kq_ch
        auto dot = [&](const float4& k) {
            return fmaf(Ab.w, k.w, fmaf(Ab.z, k.z, fmaf(Ab.y, k.y, Ab.x * k.x)));
        };
        const float Kx = dot(cc.kx), Ky = dot(cc.ky), Kz = dot(cc.kz);
        const float Kxy = dot(cc.kxy), Kxz = dot(cc.kxz), Kyz = dot(cc.kyz), Kxyz = dot(cc.kxyz);
        // d f / d t_a of the cell polynomial; t = u - corner on a uniform grid (d/du =
        // d/dt), t = (u - corner) w otherwise (d/du = w d/dt)
        rb[0] = fmaf(tz, fmaf(ty, Kxyz, Kxz), fmaf(ty, Kxy, Kx));
        rb[1] = fmaf(tz, fmaf(tx, Kxyz, Kyz), fmaf(tx, Kxy, Ky));
        rb[2] = fmaf(ty, fmaf(tx, Kxyz, Kyz), fmaf(tx, Kxz, Kz));
        if (!UNI) {
            rb[0] *= cc.w0;
            rb[1] *= cc.w1;
            rb[2] *= cc.w2;
        }
    }
    accumulate<FIRST>(G, out, M, in, cc.id, in ? tx : 0.f, in ? ty : 0.f, in ? tz : 0.f, Ab);
}

// Forward RHS (gradient channels) at the grid-unit point u for the lanes with a
// step (act); zero outside the box.  Collective (uniform cell reloads).
template <bool UNI>
__device__ __forceinline__ void fwd_eval(const DevGrid& G, bool act, float u0, float u1, float u2,
                                         AdjCell& cc, float4& A) {
    float tx, ty, tz;
    const bool in = locate<UNI>(G, act, u0, u1, u2, cc, tx, ty, tz);
    A = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in) poly_eval<false>(cc, tx, ty, tz, A);
}

// Transposed RK4 step.  On entry rb, vb are the adjoints of the state after the
// step; on exit, of the state before it.  tb (tau adjoint) is unchanged by a
// step.  Called by every lane together; lanes with act = false pass through.
template <bool KAPPA, bool UNI>
__device__ __forceinline__ void step_adj(const DevGrid& G, bool act, const float rh[3],
                                         const float vh[3], float h, float rb[3], float vb[3],
                                         float tb, AdjCell& cc, Moments& M,
                                         const AdjOut& out) {
    const float hh = 0.5f * h;
    const float h6 = h * (1.0f / 6.0f);
    const float h3 = h * (1.0f / 3.0f);
    const float hh6 = h * h6;
    // forward stage points (the same arithmetic as integrate())
    float4 A1 = make_float4(0.f, 0.f, 0.f, 0.f), A2 = A1;
    fwd_eval<UNI>(G, act, rh[0], rh[1], rh[2], cc, A1);
    fwd_eval<UNI>(G, act, fmaf(hh, vh[0], rh[0]), fmaf(hh, vh[1], rh[1]), fmaf(hh, vh[2], rh[2]),
                  cc, A2);
    const float r2[3] = {fmaf(hh, vh[0], rh[0]), fmaf(hh, vh[1], rh[1]), fmaf(hh, vh[2], rh[2])};
    const float v2[3] = {fmaf(hh, A1.x, vh[0]), fmaf(hh, A1.y, vh[1]), fmaf(hh, A1.z, vh[2])};
    const float r3[3] = {fmaf(hh, v2[0], rh[0]), fmaf(hh, v2[1], rh[1]), fmaf(hh, v2[2], rh[2])};
    const float v3[3] = {fmaf(hh, A2.x, vh[0]), fmaf(hh, A2.y, vh[1]), fmaf(hh, A2.z, vh[2])};
    const float r4[3] = {fmaf(h, v3[0], rh[0]), fmaf(h, v3[1], rh[1]), fmaf(h, v3[2], rh[2])};
    const float tk = KAPPA ? tb : 0.f;
    float p4[3], p3[3], p2[3], p1[3];
    // stage 4: a4 enters v' with weight h/6
    const float4 B4 = make_float4(h6 * vb[0], h6 * vb[1], h6 * vb[2], h6 * tk);
    eval_adj<true, UNI>(G, act, r4[0], r4[1], r4[2], B4, cc, M, out, p4);
    const float v3b[3] = {h * p4[0], h * p4[1], h * p4[2]};  // r4 = r + h v3
    // stage 3: a3 enters r' (h^2/6) and v' (h/3)
    const float4 B3 = make_float4(fmaf(hh6, rb[0], h3 * vb[0]), fmaf(hh6, rb[1], h3 * vb[1]),
                                  fmaf(hh6, rb[2], h3 * vb[2]), h3 * tk);
    eval_adj<false, UNI>(G, act, r3[0], r3[1], r3[2], B3, cc, M, out, p3);
    const float v2b[3] = {hh * p3[0], hh * p3[1], hh * p3[2]};  // r3 = r + h/2 v2
    // stage 2: a2 enters r' (h^2/6), v' (h/3) and v3 = v + h/2 a2
    const float4 B2 = make_float4(fmaf(hh, v3b[0], fmaf(hh6, rb[0], h3 * vb[0])),
                                  fmaf(hh, v3b[1], fmaf(hh6, rb[1], h3 * vb[1])),
                                  fmaf(hh, v3b[2], fmaf(hh6, rb[2], h3 * vb[2])), h3 * tk);
    eval_adj<false, UNI>(G, act, r2[0], r2[1], r2[2], B2, cc, M, out, p2);
    // stage 1: a1 enters r' (h^2/6), v' (h/6) and v2 = v + h/2 a1
    const float4 B1 = make_float4(fmaf(hh, v2b[0], fmaf(hh6, rb[0], h6 * vb[0])),
                                  fmaf(hh, v2b[1], fmaf(hh6, rb[1], h6 * vb[1])),
                                  fmaf(hh, v2b[2], fmaf(hh6, rb[2], h6 * vb[2])), h6 * tk);
    eval_adj<false, UNI>(G, act, rh[0], rh[1], rh[2], B1, cc, M, out, p1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // v: r2 = r + h/2 v, v2, v3, r' = r + h v + ..., v' = v + ...
        const float nv = vb[c] + h * rb[c] + v3b[c] + v2b[c] + hh * p2[c];
        const float nr = rb[c] + p1[c] + p2[c] + p3[c] + p4[c];
        rb[c] = act ? nr : rb[c];
        vb[c] = act ? nv : vb[c];
    }
}

// One ray per lane; the lanes of a wavefront walk their tapes in lockstep
// rounds so that grid-adjoint flushes can be combined across the wavefront.
// Three wavefronts per SIMD: nothing larger than a step's values is carried from
// round to round in registers (the moments go to the LDS rows, the cell
// polynomial is reloaded), so the kernel fits 168 VGPRs with a few spills, and
// 12.4 KiB of LDS per wavefront lets three 256-lane blocks share a CU.
// REPLAY = false (ckpt_every = 0): the forward pass recorded the whole
// trajectory, so the one interval's tape is read as it is, with no replay.
template <bool KAPPA, bool REPLAY, bool UNI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
adjoint_kernel(AdjArgs A) {
    const int lb = xcd_block(blockIdx.x, gridDim.x, A.nblocks);
    if (lb < 0) return;  // whole block: padding of the XCD remap
    const int64_t t = (int64_t)lb * blockDim.x + threadIdx.x;
    const bool valid = t < A.n;
    const int64_t n = A.n;
    const int64_t tt = valid ? t : 0;  // in-bounds row for the lanes past the end
    const int64_t ray = valid ? (A.order ? (int64_t)A.order[t] : t) : 0;
    // this wavefront's flush area (rows are written as 16-byte vectors)
    __shared__ __attribute__((aligned(16))) float s_flush[(256 / 64) * kWaveLds];
    const AdjOut out{A.grid_adj,
                     (SLT_LDS float*)s_flush + (threadIdx.x / 64) * kWaveLds};
    float rb[3] = {0.f, 0.f, 0.f}, vb[3] = {0.f, 0.f, 0.f}, v0[3] = {0.f, 0.f, 0.f};
    float tb = 0.f;
    if (valid) {
        // seeds are d(loss)/d(final x, v) in metres; the state is in grid units
        // (x = lo + u d, v = v_u d), so d/du = d d/dx and d/dv_u = d d/dv
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rb[c] = (float)((double)A.seed_state[c * n + ray] * A.G.dsp[c]);
            vb[c] = (float)((double)A.seed_state[(3 + c) * n + ray] * A.G.dsp[c]);
            v0[c] = velocity_units(A.G, c, A.s0[(3 + c) * n + ray]);
        }
        tb = A.seed_tau ? (float)A.seed_tau[ray] : 0.f;
    }
    AdjCell cc;
    cc.id = -1;
    cc.c0 = cc.c1 = cc.c2 = __builtin_nanf("");
    cc.face = true;
    cc.w0 = cc.w1 = cc.w2 = 1.0f;
    Moments M;
    M.id = -1;
    M.m0 = M.mx = M.my = M.mz = M.mxy = M.mxz = M.myz = M.mxyz = make_float4(0.f, 0.f, 0.f, 0.f);
    {
        // the lane's moments of its current cell live in its LDS row between rounds; row 64
        // stays zero (the member lists' filler)
        SLT_LDS slt_f4* row =
            reinterpret_cast<SLT_LDS slt_f4*>(out.lds + __lane_id() * kRowStride);
#pragma unroll
        for (int q = 0; q < 8; ++q) row[q] = slt_f4{0.f, 0.f, 0.f, 0.f};
        if (__lane_id() < kRowStride / 4)
            reinterpret_cast<SLT_LDS slt_f4*>(out.lds + 64 * kRowStride)[__lane_id()] =
                slt_f4{0.f, 0.f, 0.f, 0.f};
    }
    const SaveSink none{nullptr, nullptr, n, ray, 0, A.nseg, A.s0 + 3 * n, n, ray};
    for (int c = A.nck - 1; c >= 0; --c) {
        GlobalTape tape{A.tape, n, tt, 0};
        if (!REPLAY) tape.count = valid ? static_cast<const int32_t*>(A.ckpt)[t] : 0;
        if (REPLAY && valid) {
            const CkptView v = ckpt_view(A.ckpt, n, c);
            RayState s;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s.rb[a] = v.rb[a][t];
                s.off[a] = v.off[a][t];
                s.dv[a] = v.dv[a][t];
                s.v0[a] = v0[a];
            }
            s.tau = 0.0;
            s.fold = v.fold[t];
            s.path = 0.0;
            s.nout = 0;
            const int64_t g0 = (int64_t)c * A.K;
            const int64_t g1 = g0 + A.K < A.sched ? g0 + A.K : A.sched;
            // the tape needs the trajectory only, which kappa does not change: the
            // replay skips the kappa channel and tau
            integrate<false, false, GlobalTape, UNI>(A.G, s, A.seg_nsub, A.seg_h, A.nseg, g0,
                                                     g1, none, tape);
        }
        for (int round = 0; __ballot(round < tape.count); ++round) {
            // every lane reads a tape record (a lane whose tape is done reads its ray's
            // record of slot 0, inside the tape, and uses none of it); a straight move's
            // record holds only its duration
            const bool live = round < tape.count;
            const slt_f4* p = tape.rec(live ? tape.count - 1 - round : 0);
            const slt_f4 q0 = p[0], q1 = p[n];
            const float hT = q0.w;
            const float rh[3] = {q0.x, q0.y, q0.z}, vh[3] = {q1.x, q1.y, q1.z};
            // straight move r' = r + T v
            const float T = live && hT < 0.f ? -hT : 0.f;
#pragma unroll
            for (int a = 0; a < 3; ++a) vb[a] = fmaf(T, rb[a], vb[a]);
            const bool act = live && hT > 0.f;
            // the cell polynomial is not carried from round to round: every lane reloads its
            // cached cell (voxel 0 while it has none)
            load_cell(A.G, cc.id < 0 ? 0 : cc.id, cc);
            step_adj<KAPPA, UNI>(A.G, act, rh, vh, act ? hT : 0.f, rb, vb, tb, cc, M, out);
            // the round's sums go to the row; the next round's first stage rewrites them
            if (M.id >= 0) add_to_row(out, M);
        }
    }
    flush_wave<false>(A.G, out, M, M.id >= 0);  // every sum is in the rows by now
}

// ---------------------------------------------------------------------------
// Voxel adjoint -> dL/dn_e
// ---------------------------------------------------------------------------
struct UnpackArgs {
    const float* __restrict__ adj;  // packed layout, 4 floats per voxel
    const float* __restrict__ ne;
    const float* __restrict__ c0;
    const float* __restrict__ c1;
    const float* __restrict__ c2;
    double coeff;
    const float* __restrict__ te;
    int64_t n_te;
    const float* __restrict__ zz;
    int64_t n_z;
    double omega, e_charge, epsilon_0, c_light, coulomb_log;
    int n0, n1, n2;
    int64_t s0, s1, s2;
    double isp[3];  // the pack's per-axis scale of the gradient channels (1 / cell size)
    float* __restrict__ dne;
};

// Coefficient of f[j] in the gradient output P[i] along an axis of n points
// (numpy.gradient with coordinates, edge_order=1).
__device__ inline double grad_coef(const float* __restrict__ c, int n, int i, int j) {
    if (i == 0) {
        const double d = (double)c[1] - (double)c[0];
        return j == 0 ? -1.0 / d : (j == 1 ? 1.0 / d : 0.0);
    }
    if (i == n - 1) {
        const double d = (double)c[n - 1] - (double)c[n - 2];
        return j == n - 1 ? 1.0 / d : (j == n - 2 ? -1.0 / d : 0.0);
    }
    const double dx1 = (double)c[i] - (double)c[i - 1];
    const double dx2 = (double)c[i + 1] - (double)c[i];
    if (j == i - 1) return -dx2 / (dx1 * (dx1 + dx2));
    if (j == i) return (dx2 - dx1) / (dx1 * dx2);
    if (j == i + 1) return dx1 / (dx2 * (dx1 + dx2));
    return 0.0;
}

// d kappa / d n_e of the NRL rate (src/core/propagator.py:147-164); the
// Coulomb logarithm depends on n_e only through omega_pe when omega_pe > omega
// and it is neither clamped nor overridden.
__device__ inline double dkappa_dne(double ne, double te, double z, const UnpackArgs& A) {
    if (ne <= 0.0) return 0.0;
    const double pi = 3.14159265358979311600;
    const double ne_cc = ne * 1e-6;
    const double v_the = 4.19e5 * sqrt(te);
    const double o_pe = 5.64e4 * sqrt(ne_cc);
    const double o_max = o_pe > A.omega ? o_pe : A.omega;
    const double b_min = z * A.e_charge / (4.0 * pi * A.epsilon_0 * te);
    double cl = log(v_the / (o_max * b_min));
    double dcl = o_pe > A.omega ? -0.5 / ne : 0.0;
    if (!(cl >= 2.0)) { cl = 2.0; dcl = 0.0; }
    if (A.coulomb_log > 0.0) { cl = A.coulomb_log; dcl = 0.0; }
    const double r = ne_cc / A.omega;
    const double kap = 3.1e-5 * z * A.c_light * (r * r) * cl * pow(te, -1.5);
    return kap * (2.0 / ne + dcl / cl);
}

__global__ void __launch_bounds__(256) unpack_kernel(UnpackArgs A) {
    const int64_t total = (int64_t)A.n0 * A.n1 * A.n2;
    const int64_t sy = A.n2, sx = (int64_t)A.n1 * A.n2;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < total;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int idx[3] = {(int)(lin / sx), (int)((lin % sx) / sy), (int)(lin % sy)};
        const int64_t pk = idx[0] * A.s0 + idx[1] * A.s1 + idx[2] * A.s2;
        const int nn[3] = {A.n0, A.n1, A.n2};
        const int64_t ps[3] = {A.s0, A.s1, A.s2};
        const float* cs[3] = {A.c0, A.c1, A.c2};
        double f = 0.0;
        for (int a = 0; a < 3; ++a) {
            const int j = idx[a];
            for (int i = j - 1; i <= j + 1; ++i) {
                if (i < 0 || i >= nn[a]) continue;
                const double w = grad_coef(cs[a], nn[a], i, j);
                if (w != 0.0)
                    f += w * A.isp[a] * (double)A.adj[(pk + (int64_t)(i - j) * ps[a]) * 4 + a];
            }
        }
        double d = A.coeff * f;
        if (A.te) {
            const double te = (double)A.te[A.n_te == 1 ? 0 : lin];
            const double z = (double)A.zz[A.n_z == 1 ? 0 : lin];
            d += dkappa_dne((double)A.ne[lin], te, z, A) * (double)A.adj[pk * 4 + 3];
        }
        A.dne[lin] = (float)d;
    }
}

typedef void (*AdjKernel)(AdjArgs);

// The forward kernel for a channel count (kappa), mode (trajectory tape) and grid.
static AdjKernel forward_kernel_for(bool kap, bool traj, bool uni) {
    if (uni)
        return kap ? (traj ? forward_ckpt_kernel<true, true, true>
                           : forward_ckpt_kernel<true, false, true>)
                   : (traj ? forward_ckpt_kernel<false, true, true>
                           : forward_ckpt_kernel<false, false, true>);
    return kap ? (traj ? forward_ckpt_kernel<true, true, false>
                       : forward_ckpt_kernel<true, false, false>)
               : (traj ? forward_ckpt_kernel<false, true, false>
                       : forward_ckpt_kernel<false, false, false>);
}

// The reverse kernel for a channel count (kappa), mode (replay) and grid.
static AdjKernel adjoint_kernel_for(bool kap, bool replay, bool uni) {
    if (uni)
        return kap ? (replay ? adjoint_kernel<true, true, true> : adjoint_kernel<true, false, true>)
                   : (replay ? adjoint_kernel<false, true, true>
                             : adjoint_kernel<false, false, true>);
    return kap ? (replay ? adjoint_kernel<true, true, false> : adjoint_kernel<true, false, false>)
               : (replay ? adjoint_kernel<false, true, false> : adjoint_kernel<false, false, false>);
}

static int fill_adj_args(const void* packed, const slt_grid_desc* g, const slt_adjoint_args* a,
                         AdjArgs* A, const char* who) {
    int rc = make_dev_grid(packed, g, &A->G, /*allow_nonuniform=*/true);
    if (rc) return rc;
    if (!a || a->n_rays < 0 || a->n_seg < 1 || a->ckpt_every < 0 || a->sched_steps < 0 ||
        !a->seg_nsub || !a->seg_h || (a->n_rays > 0 && (!a->s0 || !a->checkpoints))) {
        set_error("%s: invalid arguments", who);
        return SLT_ERR_INVALID;
    }
    A->n = a->n_rays;
    A->s0 = a->s0;
    A->order = a->order;
    A->nseg = a->n_seg;
    A->K = a->ckpt_every;
    A->sched = a->sched_steps;
    // ckpt_every = 0: one interval, the whole trajectory on the forward's tape
    A->nck = a->ckpt_every ? (int)((a->sched_steps + a->ckpt_every - 1) / a->ckpt_every) : 1;
    if (A->nck < 1) A->nck = 1;
    if (a->ckpt_every == 0 && a->sched_steps > INT32_MAX) {
        set_error("%s: a trajectory tape holds at most %d moves per ray", who, INT32_MAX);
        return SLT_ERR_UNSUPPORTED;
    }
    A->seg_nsub = a->seg_nsub;
    A->seg_h = a->seg_h;
    A->ckpt = a->checkpoints;
    A->final_state = a->final_state;
    A->final_tau = a->final_tau;
    A->seed_state = a->seed_state;
    A->seed_tau = a->seed_tau;
    A->grid_adj = a->grid_adj;
    A->tape = static_cast<float*>(a->tape);
    A->total = a->total_steps;
    const int64_t nb = (a->n_rays + 255) / 256;
    if (nb > INT32_MAX / 2) {
        set_error("%s: too many rays", who);
        return SLT_ERR_UNSUPPORTED;
    }
    A->nblocks = (int)nb;
    return SLT_OK;
}

}  // namespace slt

using namespace slt;

extern "C" int64_t slt_checkpoint_bytes(int64_t n_rays, int64_t sched_steps, int32_t ckpt_every) {
    if (n_rays < 0 || sched_steps < 0 || ckpt_every < 0) return 0;
    if (ckpt_every == 0) return 4 * n_rays;  // each ray's tape record count
    int64_t nck = (sched_steps + ckpt_every - 1) / ckpt_every;
    if (nck < 1) nck = 1;
    return nck * kCkptBytesPerRay * n_rays;
}

extern "C" int64_t slt_tape_bytes(int64_t n_rays, int32_t ckpt_every) {
    if (n_rays < 0 || ckpt_every < 1) return 0;
    return (int64_t)ckpt_every * kTapeRecBytes * n_rays;
}

extern "C" int64_t slt_trajectory_tape_bytes(int64_t n_rays, int64_t sched_steps) {
    if (n_rays < 0 || sched_steps < 0 || sched_steps > INT32_MAX) return 0;
    // one record per ray and schedule step at most (a straight move covers several)
    if (n_rays > 0 && sched_steps > INT64_MAX / kTapeRecBytes / n_rays) return 0;
    return sched_steps * kTapeRecBytes * n_rays;
}

extern "C" int slt_forward_ckpt(const void* packed, const slt_grid_desc* g,
                                const slt_adjoint_args* a, slt_stream_t stream) {
    clear_error();
    AdjArgs A;
    int rc = fill_adj_args(packed, g, a, &A, "slt_forward_ckpt");
    if (rc) return rc;
    if (A.n == 0) return SLT_OK;
    if (!A.final_state) {
        set_error("slt_forward_ckpt: final_state is required");
        return SLT_ERR_INVALID;
    }
    const bool traj = a->ckpt_every == 0;
    if (traj && !A.tape) {
        set_error("slt_forward_ckpt: ckpt_every = 0 records the trajectory on a->tape "
                  "(slt_trajectory_tape_bytes bytes), which is NULL");
        return SLT_ERR_INVALID;
    }
    const int grid = xcd_pad(A.nblocks);
    hipLaunchKernelGGL(forward_kernel_for(g->nchan >= 4, traj, g->uniform != 0), dim3(grid),
                       dim3(256), 0, (hipStream_t)stream, A);
    return check_launch("slt_forward_ckpt");
}

extern "C" int slt_adjoint(const void* packed, const slt_grid_desc* g, const slt_adjoint_args* a,
                           slt_stream_t stream) {
    clear_error();
    AdjArgs A;
    int rc = fill_adj_args(packed, g, a, &A, "slt_adjoint");
    if (rc) return rc;
    if (A.n == 0) return SLT_OK;
    if (!A.seed_state || !A.grid_adj || !A.tape) {
        set_error("slt_adjoint: seed_state, grid_adj and tape are required");
        return SLT_ERR_INVALID;
    }
    const int grid = xcd_pad(A.nblocks);
    hipLaunchKernelGGL(adjoint_kernel_for(g->nchan >= 4, a->ckpt_every > 0, g->uniform != 0),
                       dim3(grid), dim3(256), 0, (hipStream_t)stream, A);
    return check_launch("slt_adjoint");
}

// ---------------------------------------------------------------------------
// Allocation-free reverse pass: the tape in a library-owned device array.
// ---------------------------------------------------------------------------

// The reverse pass's tape for slt_adjoint_noalloc, one copy per GPU (held from
// the library's load on that GPU until the process ends).  In host code this
// name is a stand-in for the symbol, not a device address: the address of each
// GPU's copy comes from hipGetSymbolAddress and is what the kernel receives.
static __device__ __align__(256) unsigned char scratch[SLT_ADJ_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the current GPU's scratch, looked up once per device.
int scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("slt_adjoint_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return SLT_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_error("slt_adjoint_noalloc: device %d beyond the %d supported", dev,
                  kMaxScratchDevices);
        return SLT_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("slt_adjoint_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return SLT_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SLT_OK;
}

}  // namespace

extern "C" int slt_adjoint_noalloc(const void* packed, const slt_grid_desc* g,
                                   const slt_adjoint_args* a, slt_stream_t stream) {
    clear_error();
    AdjArgs A;
    int rc = fill_adj_args(packed, g, a, &A, "slt_adjoint_noalloc");
    if (rc) return rc;
    if (a->tape) {
        set_error("slt_adjoint_noalloc: tape must be NULL (the tape is the library's scratch)");
        return SLT_ERR_INVALID;
    }
    if (a->ckpt_every == 0) {
        set_error("slt_adjoint_noalloc: ckpt_every = 0 needs the forward's trajectory tape; "
                  "call slt_adjoint with it");
        return SLT_ERR_UNSUPPORTED;
    }
    if (A.n == 0) return SLT_OK;
    if (!A.seed_state || !A.grid_adj) {
        set_error("slt_adjoint_noalloc: seed_state and grid_adj are required");
        return SLT_ERR_INVALID;
    }
    // the tape holds one record per ray and step of an interval: n * K <= cap, tested
    // without forming the product (it could overflow)
    const int64_t cap = SLT_ADJ_SCRATCH_BYTES / kTapeRecBytes;
    if (A.n > cap / A.K) {
        set_error("slt_adjoint_noalloc: %lld rays at ckpt_every %d exceed the %lld-byte tape "
                  "scratch (n_rays * ckpt_every <= %lld)",
                  (long long)A.n, A.K, (long long)SLT_ADJ_SCRATCH_BYTES, (long long)cap);
        return SLT_ERR_UNSUPPORTED;
    }
    void* tape = nullptr;
    rc = scratch_address(&tape);
    if (rc) return rc;
    // every tape slot is written by the replay before the reverse sweep reads it,
    // so the scratch needs no clearing between calls
    A.tape = static_cast<float*>(tape);
    const int grid = xcd_pad(A.nblocks);
    hipLaunchKernelGGL(adjoint_kernel_for(g->nchan >= 4, true, g->uniform != 0), dim3(grid),
                       dim3(256), 0, (hipStream_t)stream, A);
    return check_launch("slt_adjoint_noalloc");
}

extern "C" int slt_adjoint_unpack(const float* grid_adj, const slt_grid_desc* g, const float* ne,
                                  const float* cx, const float* cy, const float* cz,
                                  double grad_coeff, const float* te, int64_t n_te,
                                  const float* zz, int64_t n_z, double omega, double e_charge,
                                  double epsilon_0, double speed_of_light, double coulomb_log,
                                  float* dne, slt_stream_t stream) {
    clear_error();
    if (!grid_adj || !g || !ne || !cx || !cy || !cz || !dne) {
        set_error("slt_adjoint_unpack: null pointer");
        return SLT_ERR_INVALID;
    }
    const int64_t total = (int64_t)g->n[0] * g->n[1] * g->n[2];
    if ((te || zz) && (!te || !zz || (n_te != 1 && n_te != total) || (n_z != 1 && n_z != total))) {
        set_error("slt_adjoint_unpack: Te/Z must both be given with 1 or nx*ny*nz elements");
        return SLT_ERR_INVALID;
    }
    UnpackArgs A;
    A.adj = grid_adj;
    A.ne = ne;
    A.c0 = cx; A.c1 = cy; A.c2 = cz;
    A.coeff = grad_coeff;
    A.te = te; A.n_te = n_te;
    A.zz = zz; A.n_z = n_z;
    A.omega = omega; A.e_charge = e_charge; A.epsilon_0 = epsilon_0; A.c_light = speed_of_light;
    A.coulomb_log = coulomb_log;
    A.n0 = g->n[0]; A.n1 = g->n[1]; A.n2 = g->n[2];
    A.s0 = g->stride[0]; A.s1 = g->stride[1]; A.s2 = g->stride[2];
    for (int ax = 0; ax < 3; ++ax) {
        if (g->n[ax] < 2 || !(g->hi[ax] > g->lo[ax])) {
            set_error("slt_adjoint_unpack: grid axis %d needs >= 2 points and hi > lo", ax);
            return SLT_ERR_INVALID;
        }
        A.isp[ax] = grid_inv_spacing(g, ax);
    }
    A.dne = dne;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, A);
    return check_launch("slt_adjoint_unpack");
}
