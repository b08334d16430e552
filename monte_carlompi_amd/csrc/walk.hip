// walk.hip -- photon random walk in a snow layer for MI355X (gfx950).
//
// One photon per lane in a persistent grid.  Work is a queue of chunks
// (case, photon range); the waves of a block draw photons from the block's
// current chunk through an LDS counter and refill idle lanes wave-wide
// (ballot + mbcnt rank, one LDS atomic per refill).  Each chunk owns one of
// SLOTS LDS tally slots (FP64 BRF histogram 18 x 36 + sums + counts); the wave
// that finishes a chunk's last photon flushes the slot to the case's global
// tallies.  Finished photons are compacted per wave (ballot/mbcnt + one LDS
// atomic) into the chunk's contiguous id range, so record stores of lanes that
// finish together are contiguous; a second kernel puts each chunk's records in
// photon order through LDS.
//
// Two walkers share that skeleton:
//  * NativeWalker: FP32, Philox4x32-10 keyed by the seed, counter
//    (photon, case, step, stream); once the queue is empty, sparse waves
//    finish their photons in the wave-cooperative mode below and hand them
//    to waves of the block that ran out of work;
//  * ReplayWalker: FP64 with the reference's exact step semantics
//    (monte_carloMPI/monte_carlo3D.py:1212-1490), consuming the reference's
//    uniform deviates from CSR slices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "mcsnow.h"
#include "philox.h"
#include "walk_state.h"
#include "walkers.h"
#include "coop.h"

namespace mcs {

// ------------------------------------------------------------ chunk queue
// The block's current chunk: BlockState::cur = (seq << 32) | (slot << 28) |
// next photon offset (chunks are at most 2^24 photons, so hand-outs past the
// end stay below 2^28).
constexpr int CUR_SLOT_SHIFT = 28;
constexpr uint32_t CUR_OFF_MASK = (1u << CUR_SLOT_SHIFT) - 1u;

// Take the block's chunk number `seq` into free slot `slot` (whole wave calls).
// Returns false when the global queue is empty.
template <class BS>
__device__ bool acquire_chunk(BS &bs, uint32_t seq, int slot, const KParams &P)
{
    const int lane = lane_id();
    unsigned int j = 0;
    if (lane == 0) j = atomicAdd(P.chunk_counter, 1u);
    j = (unsigned int)bcast((int)j, 0);
    if (j >= P.n_chunks) {
        if (lane == 0) __atomic_store_n(&bs.exhausted, 1, __ATOMIC_RELAXED);
        return false;
    }
    // case of chunk j: last c with case_chunk0[c] <= j
    int lo = 0, hi = P.n_cases;  // invariant: chunk0[lo] <= j < chunk0[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((unsigned int)P.case_chunk0[mid] <= j) lo = mid; else hi = mid;
    }
    const DevCase &C = P.cases[lo];
    const unsigned long long p0 = (unsigned long long)(j - (unsigned int)P.case_chunk0[lo]) * (unsigned)P.chunk;
    const long long left = C.n_photons - (long long)p0;
    const int count = (int)(left < P.chunk ? left : P.chunk);
    SlotTally &T = bs.tal[slot];
    for (int w = lane; w < TF; w += 64) T.f[w] = 0.0;
    for (int w = lane; w < TU; w += 64) T.u[w] = 0ull;
    if (lane == 0) {
        SlotHdr &H = bs.hdr[slot];
        H.gid0 = (unsigned long long)C.gid_base + p0;
        H.p0 = (unsigned int)p0;
        H.case_id = lo;
        H.remaining = count;
        H.rec_next = 0;
        H.state = 1;
        __atomic_store_n(&H.seqcount, ((unsigned long long)seq << 32) | (unsigned)count,
                         __ATOMIC_RELAXED);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    return true;
}

// Flush slot s to its case's global tallies and free it (whole wave calls).
template <class BS>
__device__ void flush_slot(BS &bs, int s, const KParams &P)
{
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int lane = lane_id();
    const int cid = bs.hdr[s].case_id;
    const SlotTally &T = bs.tal[s];
    double *gf = P.tally_f + (size_t)cid * TF;
    unsigned long long *gu = P.tally_u + (size_t)cid * TU;
    for (int w = lane; w < TF; w += 64) {
        const double v = T.f[w];
        if (v != 0.0) unsafeAtomicAdd(gf + w, v);
    }
    for (int w = lane; w < TU; w += 64) {
        const unsigned long long v = T.u[w];
        if (v) atomicAdd(gu + w, v);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __atomic_store_n(&bs.hdr[s].state, 0, __ATOMIC_RELAXED);
}

// DIRECT: records go straight to photon order (pos = gid, no compaction
// buffer, no id store) -- a separate kernel instance, so the usual one keeps
// its registers
template <class W, bool DIRECT>
__device__ __forceinline__ void write_record(const KParams &P, size_t pos, uint64_t gid,
                                             const W &w, typename W::Real th,
                                             typename W::Real ph, typename W::Real pa)
{
    using R = typename W::Real;
    if constexpr (DIRECT) pos = gid;
    else P.rec_gid[pos] = (uint32_t)gid;
    P.rec_nscat[pos] = w.i - 1;
    P.rec_cb[pos] = (uint32_t)w.cond | ((uint32_t)w.band << 8);
    ((R *)P.rec_theta)[pos] = th;
    ((R *)P.rec_phi)[pos] = ph;
    ((R *)P.rec_path)[pos] = pa;
}

// Record + tally the DONE lanes of the wave; flush chunks that complete.
template <class W, bool DIRECT = false, class BS>
__device__ __forceinline__ void retire(BS &bs, const KParams &P, W &w, int &state,
                                       int my_slot, uint64_t my_gid)
{
    const int lane = lane_id();
    unsigned long long fin = __ballot(state == DONE);
    while (fin) {
        const int l0 = __ffsll((long long)fin) - 1;
        const int s = bcast(my_slot, l0);
        const unsigned long long same = __ballot(state == DONE && my_slot == s) & fin;
        fin &= ~same;
        const int cnt = __popcll(same);
        const bool mine = (same >> lane) & 1ull;
        SlotHdr &H = bs.hdr[s];
        SlotTally &T = bs.tal[s];
        int base = 0;
        if (lane == 0) base = atomicAdd(&H.rec_next, cnt);
        base = bcast(base, 0);
        if (mine) {
            typename W::Real th, ph, pa;
            bool axis;
            w.result(P, H.case_id, th, ph, pa, axis);
            if (DIRECT ? P.rec_nscat != nullptr : P.rec_gid != nullptr)
                write_record<W, DIRECT>(P, (size_t)H.gid0 + base + lane_rank(same), my_gid, w, th, ph, pa);
            if constexpr (W::STOKES) {
                if (P.rec_stokes) {   // photon order: Q, U, V planes of n_total
                    typename W::Real *st = (typename W::Real *)P.rec_stokes;
                    st[my_gid] = w.sQ;
                    st[P.n_total + my_gid] = w.sU;
                    st[2 * P.n_total + my_gid] = w.sV;
                }
            }
            const DevBand &B = P.bands[P.cases[H.case_id].band_begin + w.band];
            const double wvn = B.wvn;
            unsafeAtomicAdd(&T.f[T_QDOWN], wvn);
            atomicAdd(&T.u[U_N], 1ull);
            atomicAdd(&T.u[U_STEPS], (unsigned long long)w.i);
            if (w.cond >= 1) {
                unsafeAtomicAdd(&T.f[T_QCOND + w.cond - 1], wvn);
                atomicAdd(&T.u[w.cond], 1ull);
            } else {
                unsafeAtomicAdd(&T.f[T_QCAP], wvn);
                atomicAdd(&T.u[U_CAP], 1ull);
            }
            if (axis) atomicAdd(&T.u[U_AXIS], 1ull);
            if constexpr (W::STOKES && sizeof(typename W::Real) == 4) {
                if (w.n_rej) atomicAdd(&T.u[U_REJ], (unsigned long long)w.n_rej);
            }
            if (w.cond == 1) {
                const double thd = (double)th, phd = (double)ph;
                int jt = (int)(thd * (MCS_THETA_BINS / (PI_D / 2)));
                int kp = (int)(phd * (MCS_PHI_BINS / (2 * PI_D)));
                jt = jt < 0 ? 0 : (jt > MCS_THETA_BINS - 1 ? MCS_THETA_BINS - 1 : jt);
                kp = kp < 0 ? 0 : (kp > MCS_PHI_BINS - 1 ? MCS_PHI_BINS - 1 : kp);
                if (phd != phd) kp = 0;
                unsafeAtomicAdd(&T.f[jt * MCS_PHI_BINS + kp], wvn);
                unsafeAtomicAdd(&T.f[T_QPATH], wvn * (double)pa);
            }
            if constexpr (sizeof(typename W::Real) == 8) {
                if (w.err || w.pos != w.end) atomicAdd(P.ctl, 1ull);
            }
            state = EMPTY;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        int old = 0;
        if (lane == 0) old = atomicSub(&H.remaining, cnt);
        old = bcast(old, 0);
        if (old == cnt) flush_slot(bs, s, P);
    }
}

// Give photons of the block's current chunk to EMPTY lanes, advancing the
// block to the next chunk of the global queue when the current one is used up.
template <class W, class BS>
__device__ __forceinline__ void refill(BS &bs, const KParams &P, W &w, int &state,
                                      int &my_slot, uint64_t &my_gid)
{
    const int lane = lane_id();
    for (int attempt = 0; attempt < 2; ++attempt) {
        const unsigned long long idle = __ballot(state == EMPTY);
        const int n = __popcll(idle);
        if (n == 0) break;
        const unsigned long long cur = lds_load(&bs.cur);
        uint32_t seq = (uint32_t)(cur >> 32);
        int cslot = (int)((uint32_t)cur >> CUR_SLOT_SHIFT) & (BS::NSLOT - 1);
        uint32_t off = (uint32_t)cur & CUR_OFF_MASK;
        unsigned long long meta = lds_load(&bs.hdr[cslot].seqcount);
        uint32_t count = ((uint32_t)(meta >> 32) == seq) ? (uint32_t)meta : 0u;
        int got = 0;
        if (off < count) {
            unsigned long long old = 0;
            if (lane == 0) old = atomicAdd(&bs.cur, (unsigned long long)n);
            old = bcast64(old, 0);
            seq = (uint32_t)(old >> 32);
            cslot = (int)((uint32_t)old >> CUR_SLOT_SHIFT) & (BS::NSLOT - 1);
            off = (uint32_t)old & CUR_OFF_MASK;
            meta = lds_load(&bs.hdr[cslot].seqcount);
            count = ((uint32_t)(meta >> 32) == seq) ? (uint32_t)meta : 0u;
            got = off < count ? (int)min((uint32_t)n, count - off) : 0;
            const int rk = lane_rank(idle);
            if (state == EMPTY && rk < got) {
                const int s = cslot;
                const SlotHdr &H = bs.hdr[s];
                const uint32_t photon = H.p0 + off + rk;
                my_gid = H.gid0 + off + rk;
                my_slot = s;
                state = w.start(P, H.case_id, photon, my_gid) ? DONE : LIVE;
            }
        }
        if (got == n) break;
        // current chunk used up: one wave advances the block to the next
        int lk = 1;
        if (lane == 0) lk = atomicCAS(&bs.lock, 0, 1);
        lk = bcast(lk, 0);
        if (lk != 0) break;
        bool advanced = false;
        const unsigned long long now = lds_load(&bs.cur);
        const uint32_t seq_now = (uint32_t)(now >> 32);
        if (seq_now == seq && !lds_load(&bs.exhausted)) {
            // any free slot: a chunk whose last photons still walk holds only its own
            int ns = -1;
            for (int d = 1; d <= BS::NSLOT && ns < 0; ++d) {
                const int s = (cslot + d) & (BS::NSLOT - 1);
                if (lds_load(&bs.hdr[s].state) == 0) ns = s;
            }
            if (ns >= 0 && acquire_chunk(bs, seq_now + 1, ns, P)) {
                if (lane == 0)
                    __atomic_store_n(&bs.cur,
                                     ((unsigned long long)(seq_now + 1) << 32) | ((unsigned long long)ns << CUR_SLOT_SHIFT),
                                     __ATOMIC_RELAXED);
                advanced = true;
            } else if (ns < 0) {
                // every slot still holds a walking chunk: no acquisition can
                // tell this block that the queue ran dry, so read the counter
                // (otherwise small chunks would keep the block out of the
                // cooperative tail until its last chunk ends)
                unsigned int taken = 0u;
                if (lane == 0) taken = atomicOr(P.chunk_counter, 0u);
                taken = (unsigned int)bcast((int)taken, 0);
                if (taken >= P.n_chunks && lane == 0) __atomic_store_n(&bs.exhausted, 1, __ATOMIC_RELAXED);
            }
        } else if (seq_now != seq) {
            advanced = true;  // someone else advanced: retry the handout
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __atomic_store_n(&bs.lock, 0, __ATOMIC_RELAXED);
        if (!advanced) break;
    }
}

// Tail mode of the polarised walk.  Once the queue is empty and a wave holds
// at most 16 photons (all with the exact sampler), the wave works in rounds:
// the G = 64 / 2^ceil(log2 L) lanes of group g build the packets of the next
// G steps of the group's photon (Philox blocks, cos theta, phase matrix,
// the phi candidates: everything that does not depend on the photon's
// state), then each photon's own lane walks those G steps through
// polar_move.  The serial chain of a step keeps only the phi acceptance,
// rotation, Stokes update and events.  Records equal the one-lane walk's.
// LDS after the tables: 16 owner lanes and PK_SLOTS packets per wave (64 and
// one that the last group's prefetch may read past its own).
constexpr int PK_SLOTS = 65;
__device__ void polar_tail(PolarWalker &w, int &state, const KParams &P, int wave)
{
    extern __shared__ __align__(16) float pm_lds_tab[];
    int *const owners = reinterpret_cast<int *>(pm_lds_tab + P.pm_lds_floats);
    // packets 16-byte aligned after the tables and owner lanes (the launch
    // reserves 16 bytes for this)
    const int pk_off = (P.pm_lds_floats + (int)(blockDim.x >> 6) * 16 + 3) & ~3;
    PolarPacket *const pk = reinterpret_cast<PolarPacket *>(pm_lds_tab + pk_off) + wave * PK_SLOTS;
    int *const own = owners + wave * 16;
    const int lane = lane_id();
    for (;;) {
        const unsigned long long live = __ballot(state == LIVE);
        if (live == 0ull) break;
        const int L = __popcll(live);
        const int lg = (L <= 1) ? 0 : 32 - __clz(L - 1);
        const int G = 64 >> lg;
        if (state == LIVE) own[lane_rank(live)] = lane;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int g = lane >> (6 - lg);
        const int o = own[g < L ? g : 0];
        const uint32_t ck0 = bperm(o, w.pre.ck0), h0 = bperm(o, w.pre.h0), l0 = bperm(o, w.pre.l0);
        const uint32_t io = bperm(o, w.i);
        const int pmo = (int)bperm(o, (uint32_t)w.pm);
        const float mass = bperm(o, w.pm_mass);
        if (g < L) {
            const PhiloxPre pre{ck0, h0, l0};
            const uint32_t st = io + 1u + (uint32_t)(lane & (G - 1));
            const u32x4 mainb = philox_step(pre, st, P.key0, P.key1);
            const u32x4 pb = philox4x32_10(u32x4{l0 * 0x991A7CDBu, ck0 ^ P.key0, st, 2u}, P.key0, P.key1);
            const float *T = (P.pm_lds_floats > 0 ? pm_lds_tab : P.pmf) + (size_t)pmo * (size_t)P.pm_stride;
            PolarPacket q;
            float r2;
            int seg;
            PolarWalker::make_packet(mainb, pb, T, P.pm_n, mass, q, r2, seg);
            pk[lane] = q;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (state == LIVE) {
            // the moves run straight through the group's packets, all photons
            // of the wave in step: the round ends for every lane after the
            // step in which any photon has an event (rare) or reaches the step
            // cap, or after the last packet -- a wave-uniform exit, so the
            // loop issues no exec-mask changes.  A photon without an event
            // simply goes on in the next round, whose packets are built from
            // its step count then; after an event (a bottom re-emission walks
            // a step of its own) likewise.  The next step's packet is loaded
            // while this one's step runs (two packets in registers, their
            // roles alternating; the last prefetch may read the slot after the
            // group's packets, never used).  Photons in the tail have i <
            // max_steps, so the first step is always taken.
            const PolarPacket *const mine = pk + lane_rank(live) * G;
            // the step cap in a VGPR for the loop (as a kernel argument it
            // would be reloaded from spilled scalar registers every step)
            uint32_t cap = P.max_steps;
            asm volatile("" : "+v"(cap));
            PolarPacket a = mine[0];
            int k = 0;
            for (;;) {
                PolarPacket b = mine[k + 1];
                bool ev = w.polar_move(P, a);
                ++k;
                if (__ballot(ev || w.i >= cap) != 0ull || k >= G) break;
                a = mine[k + 1];
                ev = w.polar_move(P, b);
                ++k;
                if (__ballot(ev || w.i >= cap) != 0ull || k >= G) break;
            }
            // the last step's event test, formed again from the state it left
            // (as polar_move forms it), so the loop carries no per-lane flag
            const PolarPacket &last = mine[k - 1];
            const bool ev = w.z > 0.f || w.z < -w.tau_tot || last.w < w.t_abs_all;
            const bool d = ev && w.polar_event(P, last.dtau, last.w);
            if (d || w.i >= P.max_steps) {
                if (!d) w.cond = 0;
                state = DONE;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// One step of the hot loop (the native walker's step form by build)
template <int WPE, class W>
__device__ __forceinline__ bool hot_walk_step(W &w, const KParams &P)
{
    if constexpr (W::COOP) return w.template step<(WPE <= 4)>(P);
    else return w.step(P);
}

// Tail mode of the reproducible HG walk (coop_live = 0).  Once the queue is
// empty and a wave holds at most HG_TAIL photons, the wave works in rounds as
// the polarised tail does: the G = HG_PK / 2^ceil(log2 L) lanes of group g
// build the packets of the next G steps of the group's photon (Philox block,
// free path, cos theta, the azimuth's sine and cosine: hg_packet), then every
// photon's own lane walks those G steps through hg_move -- rotation, move and
// event test, ~25 instructions instead of the whole step -- until the step in
// which any photon of the wave has an event or reaches the step cap, or the
// last packet (a wave-uniform exit).  The packet and the move use the
// expressions of NativeWalkerT::step and the event its hg_event, so records
// equal the one-lane walk's bit for bit.  LDS (the dynamic area, which holds
// no cooperative tables at coop_live = 0): HG_TAIL owner lanes and
// HG_PK_SLOTS packets per wave.
constexpr int HG_TAIL = 16;
constexpr int HG_PK_LOG = 6;
constexpr int HG_PK = 1 << HG_PK_LOG;
constexpr int HG_PK_SLOTS = HG_PK + 1;   // the last group's prefetch may read one past
template <int WPE>
constexpr bool HG_TAIL_IN = WPE <= 4;    // the builds that carry the packet tail
// live photons at which a wave with the queue dry leaves the hot loop for its
// tail: the cooperative tail's threshold, or the packet tail's at coop_live 0
template <int WPE>
__device__ __forceinline__ int tail_live(const KParams &P)
{
    if constexpr (HG_TAIL_IN<WPE>) return P.coop_live > 0 ? P.coop_live : HG_TAIL;
    else return P.coop_live;
}
// (a template, so that only the builds that carry it instantiate it: another
// caller of the walker's non-inlined helpers changes how the others inline)
template <int WPE>
__device__ void hg_tail(NativeWalker &w, int &state, const KParams &P, int wave)
{
    extern __shared__ __align__(16) CoopEntry coop_tab[];
    int *const owners = reinterpret_cast<int *>(coop_tab);
    HgPacket *const pk = reinterpret_cast<HgPacket *>(owners + (blockDim.x >> 6) * HG_TAIL) + wave * HG_PK_SLOTS;
    int *const own = owners + wave * HG_TAIL;
    const int lane = lane_id();
    for (;;) {
        const unsigned long long live = __ballot(state == LIVE);
        if (live == 0ull) break;
        const int L = __popcll(live);
        const int lg = (L <= 1) ? 0 : 32 - __clz(L - 1);
        const int G = HG_PK >> lg;                 // >= 2 (L <= HG_TAIL)
        if (state == LIVE) own[lane_rank(live)] = lane;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int g = lane >> (HG_PK_LOG - lg);    // >= L for lanes >= HG_PK
        const int o = own[g < L ? g : 0];
        const uint32_t ck0 = bperm(o, w.pre.ck0), h0 = bperm(o, w.pre.h0), l0 = bperm(o, w.pre.l0);
        const uint32_t io = bperm(o, w.i);
        const float ca = bperm(o, w.a), cb = bperm(o, w.b), ccc = bperm(o, w.cc), cd = bperm(o, w.d), ce = bperm(o, w.e);
        const bool iso = bperm(o, (uint32_t)w.iso) != 0u;
        if (g < L) {
            const PhiloxPre pre{ck0, h0, l0};
            const u32x4 cur = philox_step(pre, io + 1u + (uint32_t)(lane & (G - 1)), P.key0, P.key1);
            HgPacket q;
            NativeWalker::hg_packet(cur, ca, cb, ccc, cd, ce, iso, q);
            pk[lane] = q;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (state == LIVE) {
            // photons in the tail have i < max_steps: the first step is taken
            const HgPacket *const mine = pk + lane_rank(live) * G;
            uint32_t cap = P.max_steps;
            asm volatile("" : "+v"(cap));
            HgPacket a = mine[0];
            int k = 0;
            for (;;) {
                const HgPacket b = mine[k + 1];
                bool ev = w.hg_move(a);
                ++k;
                if (__ballot(ev || w.i >= cap) != 0ull || k >= G) break;
                a = mine[k + 1];
                ev = w.hg_move(b);
                ++k;
                if (__ballot(ev || w.i >= cap) != 0ull || k >= G) break;
            }
            // the last step's event test, formed again from the state it left
            const HgPacket &last = mine[k - 1];
            const bool ev = w.z > 0.f || w.z < -w.tau_tot || last.w < w.t_abs_all;
            bool d = false;
            if (ev) {
                w.r = philox_step(w.pre, w.i + 1u, P.key0, P.key1);   // the next step's block (re-emission)
                d = w.hg_event(P, last.dtau, last.w);
            }
            if (d || w.i >= P.max_steps) {
                if (!d) w.cond = 0;
                state = DONE;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// Wave issue priority from the oldest live photon (i = its step count, 0 for
// other lanes): long walks set the tail of a launch, so their waves win issue
// slots over young ones.
__device__ __forceinline__ void set_priority(uint32_t i)
{
    if (__ballot(i > 6000u)) __builtin_amdgcn_s_setprio(3);
    else if (__ballot(i > 3000u)) __builtin_amdgcn_s_setprio(2);
    else if (__ballot(i > 1500u)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// WPE: waves per SIMD the register allocation targets (the launch picks the
// instance that fits its shape: 4 -> 128 VGPRs, no spills; 5 -> 96; 6 -> 80).
// The hand-over of cooperative-tail photons between waves is left out of the
// 6-wave build: at 80 VGPRs its code would spill in the step loop.  SL: chunk
// slots of the block (tally slots in LDS, 5.3 KB each); the polarised walk
// in blocks of up to 256 threads takes 2, so that with its tables and tail
// packets three blocks fit in a CU's LDS (with 4, two do; larger blocks keep
// 4, or their waves would wait for slots held by lingering walks).
template <class W, int WPE, bool DIRECT = false, int SL = SLOTS>
__global__ void __launch_bounds__(W::MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_walk(const KParams P)
{
    constexpr bool HAND = W::COOP && WPE <= 5;
    __shared__ BlockStateT<SL> bs;
    // native mode: coop_live table entries per wave, then coop_live group slots per wave
    extern __shared__ CoopEntry coop_tab[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    // ---- init: zero LDS, wave 0 takes the first chunk
    {
        unsigned long long *w64 = reinterpret_cast<unsigned long long *>(&bs);
        for (int k = threadIdx.x; k < (int)(sizeof(bs) / 8); k += blockDim.x) w64[k] = 0ull;
    }
    if constexpr (W::STOKES && W::COOP == false && sizeof(typename W::Real) == 4) {
        // polarised native walk: phase-matrix tables into dynamic LDS
        if (P.pm_lds_floats > 0) {
            extern __shared__ __align__(16) float pm_lds_tab[];
            for (int k = threadIdx.x; k < P.pm_lds_floats; k += blockDim.x) pm_lds_tab[k] = P.pmf[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) bs.holders = blockDim.x >> 6;
    if (wave == 0) {
        if (!acquire_chunk(bs, 0u, 0, P)) {
            if (lane == 0) bs.hdr[0].seqcount = 0ull;  // seq 0, count 0
        }
    }
    __syncthreads();

    int *const coop_map = reinterpret_cast<int *>(coop_tab + (blockDim.x >> 6) * P.coop_live);
    W w;
    int state = EMPTY;
    int my_slot = 0;
    uint64_t my_gid = 0;
    int spins = 0;
    bool to_coop = false;

    for (;;) {
        // ---------------- cold: retire finished lanes, refill empty ones
        if (__ballot(state == DONE)) retire<W, DIRECT>(bs, P, w, state, my_slot, my_gid);
        const bool exhausted = lds_load(&bs.exhausted) != 0;
        if (!exhausted && __ballot(state == EMPTY)) refill(bs, P, w, state, my_slot, my_gid);
        const unsigned long long live = __ballot(state == LIVE);
        if (live == 0ull) {
            // photons can finish at hand-out (first step, surface mode): retire first
            if (__ballot(state == DONE)) continue;
            if (lds_load(&bs.exhausted)) break;
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 28)) break;  // cannot happen: chunks always finish
            continue;
        }
        if constexpr (W::COOP) {
            // queue empty and few photons left: walk them cooperatively
            if (lds_load(&bs.exhausted) && __popcll(live) <= P.coop_live) {
                if constexpr (HAND) {
                    to_coop = true;   // below, with the hand-over
                } else {
                    __builtin_amdgcn_s_setprio(3);
                    coop_phase<false, coop_k<WPE>()>(bs, w, state, P, coop_tab + wave * P.coop_live, coop_map + wave * P.coop_live,
                                      &bs.tal[my_slot].u[U_COOP], &bs.tal[my_slot].u[U_MOVED], wave);
                    retire<W, DIRECT>(bs, P, w, state, my_slot, my_gid);
                }
                break;
            }
            // reproducible mode (no cooperative tail): the packet tail mode,
            // in the 128-VGPR build (in the smaller ones its code would add
            // spills to the step loop; they keep the one-lane drain)
            if constexpr (HG_TAIL_IN<WPE>) {
                if (P.coop_live == 0 && lds_load(&bs.exhausted) && __popcll(live) <= HG_TAIL) {
                    __builtin_amdgcn_s_setprio(3);
                    hg_tail<WPE>(w, state, P, wave);
                    retire<W, DIRECT>(bs, P, w, state, my_slot, my_gid);
                    break;
                }
            }
        }
        constexpr bool PTAIL = W::STOKES && !W::COOP && sizeof(typename W::Real) == 4;
        const int ptail = PTAIL ? min(P.coop_live, 16) : 0;
        if constexpr (PTAIL) {
            // polarised walk: queue empty, few photons, all with the exact sampler
            if (ptail > 0 && lds_load(&bs.exhausted) && __popcll(live) <= ptail &&
                __ballot(state == LIVE && !(w.pm >= 0 && w.pm_exact)) == 0ull) {
                __builtin_amdgcn_s_setprio(3);
                polar_tail(w, state, P, wave);
                retire<W, DIRECT>(bs, P, w, state, my_slot, my_gid);
                break;
            }
        }
        // leave the hot loop once `thresh` lanes are not live (with the queue
        // empty: once the cooperative tail can take over)
        const int thresh = lds_load(&bs.exhausted) ? (W::COOP ? 64 - tail_live<WPE>(P) : 64 - ptail)
                                                   : min(64, (64 - __popcll(live)) + (int)P.refill_min);
        // ---------------- hot: step every live photon
        // Two steps per loop turn, the exit test after the second: the next
        // step's Philox block, the direction and the step count then
        // alternate between two register sets instead of being copied back
        // at the end of every turn.  Each photon takes the same steps (a lane
        // that ends in the first stays DONE).  Not in the 96-VGPR build,
        // where it would add a spill reload to the step loop.
        constexpr bool TWO = !W::STOKES && WPE != 5;
        for (uint32_t iter = 1;; ++iter) {
            if constexpr (TWO) {
                if (state == LIVE) {
                    const bool d = hot_walk_step<WPE>(w, P);
                    if (d || (!W::CAP_IN_STEP && w.i >= P.max_steps)) {
                        if (!d) w.cond = 0;
                        state = DONE;
                    }
                }
            }
            if (state == LIVE) {
                const bool d = hot_walk_step<WPE>(w, P);
                if (d || (!W::CAP_IN_STEP && w.i >= P.max_steps)) {
                    if (!d) w.cond = 0;
                    state = DONE;
                }
            }
            const unsigned long long nl = __ballot(state == LIVE);
            if (nl == 0ull || 64 - __popcll(nl) >= thresh) break;
            if ((iter & (TWO ? 127u : 255u)) == 0u) set_priority(state == LIVE ? w.i : 0u);
        }
    }
    if constexpr (HAND) {
        // the cooperative tail: the wave's last photons, then photons taken
        // over from siblings (at most coop_live / 2 at a time) until the block
        // has none left
        if (P.coop_live > 0) {
            for (;;) {
                if (to_coop) {
                    __builtin_amdgcn_s_setprio(3);
                    await_own_posts(bs, P, wave);
                    coop_phase<true, coop_k<WPE>()>(bs, w, state, P, coop_tab + wave * P.coop_live, coop_map + wave * P.coop_live,
                                     &bs.tal[my_slot].u[U_COOP], &bs.tal[my_slot].u[U_MOVED], wave);
                    // every photon the wave held is finished or handed over
                    retire<W, DIRECT>(bs, P, w, state, my_slot, my_gid);
                }
                if (!coop_receive(bs, P, w, state, my_slot, my_gid, coop_tab, wave)) break;
                to_coop = true;
            }
        }
    }
}

// records: compacted order -> photon order
template <class R>
__global__ void k_scatter_records(const uint32_t *gid, const uint32_t *nscat, const uint32_t *cb,
                                  const R *th, const R *ph, const R *pa, uint64_t n,
                                  uint32_t *o_nscat, uint32_t *o_cb, R *o_th, R *o_ph, R *o_pa)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = gid[j];
        if (g >= n) continue;
        o_nscat[g] = nscat[j];
        o_cb[g] = cb[j];
        o_th[g] = th[j];
        o_ph[g] = ph[j];
        o_pa[g] = pa[j];
    }
}

// records: compacted order -> photon order, one chunk per block pass.  A
// chunk's records sit at its own id range [gid0, gid0 + count) in any order
// (retirement order), so the block gathers them into LDS by id and writes
// them out contiguously: coalesced reads and writes.  LDS: chunk x (8 + 3
// sizeof(R)) bytes.
// Host-layout record columns (mcs_records), written on the device so that
// read-back is plain copies.  Every pointer is nullable.
struct WideRecords {
    int8_t *condition;
    uint8_t *band;
    int64_t *n_scat;
    double *theta_n, *phi_n, *path_length, *wvn, *snow_depth;
};

// Photon-order compact records (n_scat, condition | band << 8, theta, phi,
// path in R) -> WideRecords.  The case of photon j is found by bisection on
// the cases' first global ids; wvn and snow_depth come from the photon's band.
template <class R>
__global__ void k_widen_records(const DevCase *cases, const DevBand *bands, int n_cases, uint64_t n,
                                const uint32_t *nscat, const uint32_t *cb, const R *th, const R *ph, const R *pa,
                                WideRecords o)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = cb[j];
        if (o.condition) o.condition[j] = (int8_t)(v & 0xffu);
        if (o.band) o.band[j] = (uint8_t)(v >> 8);
        if (o.n_scat) o.n_scat[j] = (int64_t)nscat[j];
        if (o.theta_n) o.theta_n[j] = (double)th[j];
        if (o.phi_n) o.phi_n[j] = (double)ph[j];
        if (o.path_length) o.path_length[j] = (double)pa[j];
        if (o.wvn || o.snow_depth) {
            int lo = 0, hi = n_cases;   // cases[lo].gid_base <= j < cases[hi].gid_base
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if ((uint64_t)cases[mid].gid_base <= j) lo = mid; else hi = mid;
            }
            const DevBand &B = bands[cases[lo].band_begin + (int)(v >> 8)];
            if (o.wvn) o.wvn[j] = B.wvn;
            if (o.snow_depth) o.snow_depth[j] = B.snow_depth;
        }
    }
}

// Stokes planes (3 x n in R) -> float64 planes.
template <class R>
__global__ void k_widen_stokes(const R *in, uint64_t n3, double *out)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n3; j += (uint64_t)gridDim.x * blockDim.x)
        out[j] = (double)in[j];
}

template <class R>
__global__ void k_order_records(const DevCase *cases, const int32_t *chunk0, int n_cases, int chunk,
                                uint32_t n_chunks, const uint32_t *gid, const uint32_t *nscat, const uint32_t *cb,
                                const R *th, const R *ph, const R *pa, uint32_t *o_nscat, uint32_t *o_cb, R *o_th,
                                R *o_ph, R *o_pa)
{
    extern __shared__ uint32_t ord_lds[];
    uint32_t *s_ns = ord_lds, *s_cb = ord_lds + chunk;
    R *s_th = reinterpret_cast<R *>(ord_lds + 2 * chunk), *s_ph = s_th + chunk, *s_pa = s_ph + chunk;
    for (uint32_t j = blockIdx.x; j < n_chunks; j += gridDim.x) {
        int lo = 0, hi = n_cases;   // chunk0[lo] <= j < chunk0[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)chunk0[mid] <= j) lo = mid; else hi = mid;
        }
        const DevCase &C = cases[lo];
        const int64_t p0 = (int64_t)(j - (uint32_t)chunk0[lo]) * chunk;
        const int count = (int)min<int64_t>(chunk, C.n_photons - p0);
        const uint64_t g0 = (uint64_t)(C.gid_base + p0);
        for (int t = threadIdx.x; t < count; t += blockDim.x) {
            const uint64_t pos = g0 + t;
            const uint32_t k = gid[pos] - (uint32_t)g0;
            if (k < (uint32_t)count) {
                s_ns[k] = nscat[pos];
                s_cb[k] = cb[pos];
                s_th[k] = th[pos];
                s_ph[k] = ph[pos];
                s_pa[k] = pa[pos];
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < count; t += blockDim.x) {
            const uint64_t pos = g0 + t;
            o_nscat[pos] = s_ns[t];
            o_cb[pos] = s_cb[t];
            o_th[pos] = s_th[t];
            o_ph[pos] = s_ph[t];
            o_pa[pos] = s_pa[t];
        }
        __syncthreads();
    }
}

// Zero the tallies, counts and control words of a walk (one launch; a kernel,
// not a memset, so that the walk can be captured in a graph).
__global__ void k_clear(unsigned long long *a, uint64_t na, unsigned long long *b, uint64_t nb,
                        unsigned long long *c, uint64_t nc)
{
    const uint64_t n = na + nb + nc;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        if (j < na) a[j] = 0ull;
        else if (j < na + nb) b[j - na] = 0ull;
        else c[j - na - nb] = 0ull;
    }
}

// Chunk table of a walk on the device: chunk0[k] = sum over cases j < k of
// ceil(n_photons_j / chunk), k = 0..n_cases (one block of 1024 threads; each
// thread sums a contiguous run of cases, then a scan over the threads).
__global__ void __launch_bounds__(1024) k_chunk_table(const DevCase *cases, int n_cases, int chunk, int32_t *chunk0)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (n_cases + 1023) / 1024;
    const int k0 = min(t * per, n_cases), k1 = min(k0 + per, n_cases);
    int64_t s = 0;
    for (int k = k0; k < k1; ++k) s += (cases[k].n_photons + chunk - 1) / chunk;
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive scan (Hillis-Steele)
        const int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;   // chunks before this thread's first case
    for (int k = k0; k < k1; ++k) {
        chunk0[k] = (int32_t)run;
        run += (cases[k].n_photons + chunk - 1) / chunk;
    }
    if (t == 1023) chunk0[n_cases] = (int32_t)part[1023];
}

// The walk's checks for the host: photons, photon-steps, capped, phi on an
// axis (summed over cases), control words 0 and 1, and a written-flag, into
// pinned host memory (one block of 256 threads).
__global__ void k_summary(const unsigned long long *tu, int nc, const unsigned long long *ctl,
                          unsigned long long *out)
{
    __shared__ unsigned long long red[4][256];
    const int t = threadIdx.x;
    unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int k = t; k < nc; k += 256) {
        const unsigned long long *u = tu + (size_t)k * TU;
        a0 += u[U_N]; a1 += u[U_STEPS]; a2 += u[U_CAP]; a3 += u[U_AXIS];
    }
    red[0][t] = a0; red[1][t] = a1; red[2][t] = a2; red[3][t] = a3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + h];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0][0]; out[1] = red[1][0]; out[2] = red[2][0]; out[3] = red[3][0];
        out[4] = ctl[0]; out[5] = ctl[1]; out[6] = ctl[2];
        out[7] = 1ull;
    }
}

// Scratch of mcs_walk_noalloc, one copy per device for the life of the
// process: 16 walk control words (ctl[0..7], the chunk counter at word 8)
// then the chunk table (MCS_NOALLOC_MAX_CASES + 1 int32).  Host code uses
// the address hipGetSymbolAddress gives for the current device (noalloc_base),
// never the array's name.
constexpr size_t NOALLOC_CTL_U64 = 16;
constexpr size_t NOALLOC_SCRATCH_U64 = NOALLOC_CTL_U64 + (MCS_NOALLOC_MAX_CASES + 1 + 1) / 2;
__device__ __align__(256) unsigned long long scratch[NOALLOC_SCRATCH_U64];

__global__ void k_philox(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out, uint64_t n)
{
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const u32x4 r = philox4x32_10(u32x4{ctr[4 * j], ctr[4 * j + 1], ctr[4 * j + 2], ctr[4 * j + 3]}, k0, k1);
    out[4 * j] = r.x; out[4 * j + 1] = r.y; out[4 * j + 2] = r.z; out[4 * j + 3] = r.w;
}

}  // namespace mcs

// =================================================================== host
using namespace mcs;

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct mcs_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char err[512] = {0};
    mcs_launch launch{256, 4, 2048, 8, 16};
    // host copies
    std::vector<mcs_case> cases;
    std::vector<mcs_band> bands;
    std::vector<DevCase> dcases;
    std::vector<DevBand> dbands;
    int64_t total_photons = 0;
    bool cases_set = false;
    // device buffers
    DevBuf d_cases, d_bands, d_chunk0, d_tally_f, d_tally_u, d_ctl, d_pband, d_rdev, d_roff;
    DevBuf d_rec[6], d_out[5];
    // aspherical particles: phase-matrix tables (FP64, FP32), final Stokes
    bool polar = false;
    int32_t pm_n = 0;
    size_t pm_floats = 0;        // FP32 table floats (n_tables x pm_stride)
    DevBuf d_pm, d_pmf, d_stokes;
    DevBuf d_wide;                  // host-layout record columns for read-back
    bool last_stokes = false;
    int last_mode = MCS_NATIVE;
    bool last_records = false;
    int64_t last_n = 0;
    bool pending = false;
    mcs_stats stats{};
    // graph cache
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t gkey = 0;
    // chunk table on the device (uploaded when it changes)
    std::vector<int32_t> chunk0;
    bool chunk0_valid = false;
    // pinned host words written by k_summary (h_sum) and its device address
    unsigned long long *h_sum = nullptr, *d_sum = nullptr;
    // the last walk's tallies and counts are in the context's buffers (false
    // after mcs_walk_noalloc, which writes them to the caller's)
    bool last_tallies_ctx = true;
    // pinned host staging of record read-back (grown on demand)
    void *h_stage = nullptr;
    size_t stage_cap = 0;
    // device address of the mcs_walk_noalloc scratch on this context's device
    unsigned long long *noalloc_scratch = nullptr;
    hipError_t noalloc_err = hipSuccess;
};

// Device address of `scratch` per device, looked up once per device (with
// that device current) by the first mcs_create on it.
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch_addr[64];

static hipError_t noalloc_base(int device, unsigned long long **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = (unsigned long long *)p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

static int set_err(mcs_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_err(ctx, e_ == hipErrorOutOfMemory ? MCS_ERR_NOMEM : MCS_ERR_HIP, \
                           "%s: %s", #expr, hipGetErrorString(e_));                    \
    } while (0)

// Allocation-failure injection for tests of the MCS_ERR_NOMEM paths: with
// MCS_FAIL_ALLOC_BYTES=N in the environment (read at every allocation), any
// device buffer of N bytes or more that the library would allocate fails as
// if the device were out of memory.
static bool injected_alloc_failure(size_t bytes)
{
    const char *e = std::getenv("MCS_FAIL_ALLOC_BYTES");
    if (!e || !*e) return false;
    char *end = nullptr;
    const unsigned long long n = std::strtoull(e, &end, 10);
    return end != e && bytes >= n;
}

static int ensure(mcs_ctx *c, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return MCS_OK;
    if (b.p) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    c->gkey = 0;  // buffer addresses change: recapture any graph
    if (injected_alloc_failure(bytes))
        return set_err(c, MCS_ERR_NOMEM, "hipMalloc(%zu): out of memory (injected by MCS_FAIL_ALLOC_BYTES)", bytes);
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess)
        return set_err(c, MCS_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    b.cap = bytes;
    c->gkey = 0;  // buffer addresses changed: recapture any graph
    return MCS_OK;
}

// Walker/Vose alias table over a case's bands.  Band k's probability is what
// "first k with u < cdf_k" gives (cdf clamped to [0, 1], mass beyond the
// running maximum; the last band takes the rest), so any cdf column keeps its
// meaning.
static void build_alias(DevBand *db, const mcs_band *b, int n)
{
    std::vector<double> q(n);
    double m = 0.0;
    for (int k = 0; k < n - 1; ++k) {
        const double ck = std::min(std::max(b[k].cdf, 0.0), 1.0);
        q[k] = std::max(ck - m, 0.0) * n;
        m = std::max(m, ck);
    }
    q[n - 1] = (1.0 - m) * n;
    std::vector<int> small, large;
    for (int k = 0; k < n; ++k) (q[k] < 1.0 ? small : large).push_back(k);
    for (int k = 0; k < n; ++k) { db[k].alias_k = k; db[k].alias_t = 0xFFFFFFFFu; }
    while (!small.empty() && !large.empty()) {
        const int s = small.back(), l = large.back();
        small.pop_back();
        db[s].alias_t = (uint32_t)std::min(q[s] * 4294967296.0, 4294967295.0);
        db[s].alias_k = l;
        q[l] -= 1.0 - q[s];
        if (q[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
}

// Does the scattering matrix [[P11, P12, 0, 0], [P12, P22, 0, 0], [0, 0, P33,
// -P43], [0, 0, P43, P44]] map every fully polarised vector (1, q, u, v),
// q^2 + u^2 + v^2 = 1, to a vector of degree of polarisation <= 1?  With r^2 =
// u^2 + v^2 = 1 - q^2 the (U, V) part reaches at most s r, s the largest
// singular value of [[P33, -P43], [P43, P44]], so the condition is
//   f(q) = (P11 + P12 q)^2 - (P12 + P22 q)^2 - s^2 (1 - q^2) >= 0 on [-1, 1],
// a quadratic in q: its ends and, if inside, its vertex.  Tolerance: the
// table files carry ~7 significant digits.
static bool keeps_physical(double p11, double p12, double p22, double p33, double p43, double p44)
{
    const double ss = p33 * p33 + 2.0 * p43 * p43 + p44 * p44;
    const double det = p33 * p44 + p43 * p43;
    const double s2 = 0.5 * (ss + std::sqrt(std::max(ss * ss - 4.0 * det * det, 0.0)));
    const double A = p12 * p12 - p22 * p22 + s2;
    const double B = 2.0 * p12 * (p11 - p22);
    const double C = p11 * p11 - p12 * p12 - s2;
    const double tol = -1e-5 * p11 * p11;
    if (A + B + C < tol || A - B + C < tol) return false;
    if (A > 0.0) {
        const double qv = -B / (2.0 * A);
        if (qv > -1.0 && qv < 1.0 && C - B * B / (4.0 * A) < tol) return false;
    }
    return true;
}

static uint32_t prob_bits(double p)
{
    double t = p * 4294967296.0;
    if (!(t > 0.0)) return 0u;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)llround(t);
}

// Pinned host staging of at least `bytes` (read-back copies run at full
// PCIe rate into it, then host threads widen into the caller's arrays).
static int stage(mcs_ctx *c, size_t bytes)
{
    if (c->stage_cap >= bytes) return MCS_OK;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->stage_cap = 0;
    HIPCHK(c, hipHostMalloc(&c->h_stage, bytes, hipHostMallocDefault));
    c->stage_cap = bytes;
    return MCS_OK;
}

// Run f(j0, j1) over [0, n) on up to 16 host threads (slices of >= 64 Ki).
template <class F>
static void parallel_rows(int64_t n, F f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t T = std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::min(hw, 16u), (n + 65535) / 65536}));
    if (T == 1) { f((int64_t)0, n); return; }
    std::vector<std::thread> th;
    const int64_t per = (n + T - 1) / T;
    for (int64_t t = 1; t < T; ++t) {
        const int64_t a = std::min(n, t * per), b = std::min(n, a + per);
        th.emplace_back([=] { f(a, b); });
    }
    f(0, std::min(n, per));
    for (auto &x : th) x.join();
}

// [p, p + bytes) lies in one pinned (page-locked, hipHostMalloc) host
// allocation: the DMA engine can write it directly.
static bool host_pinned(const void *p, size_t bytes)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // pageable memory: the query reports invalid value
        return false;
    }
    if (a.type != hipMemoryTypeHost) return false;
    void *start = nullptr;
    size_t size = 0;
    if (hipPointerGetAttribute(&start, HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR, (hipDeviceptr_t)p) != hipSuccess ||
        hipPointerGetAttribute(&size, HIP_POINTER_ATTRIBUTE_RANGE_SIZE, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return (const char *)p >= (const char *)start && (const char *)p + bytes <= (const char *)start + size;
}

// Copy k device columns (d_wide + off[k], n * esz[k] bytes) to the host
// arrays dst[k] (nullable): pinned destinations by one DMA each, the others
// into pinned staging and from there by host threads.  Waits for the stream.
static int copy_columns(mcs_ctx *c, void *const *dst, const size_t *esz, const size_t *off, int k, size_t n)
{
    bool pinned[8] = {};
    size_t soff[8] = {}, stot = 0;
    for (int i = 0; i < k; ++i) {
        if (!dst[i]) continue;
        pinned[i] = host_pinned(dst[i], esz[i] * n);
        if (!pinned[i]) { soff[i] = stot; stot += esz[i] * n; }
    }
    int r;
    if (stot && (r = stage(c, stot))) return r;
    const char *w = (const char *)c->d_wide.p;
    for (int i = 0; i < k; ++i) {
        if (!dst[i]) continue;
        void *to = pinned[i] ? dst[i] : (void *)((char *)c->h_stage + soff[i]);
        HIPCHK(c, hipMemcpyAsync(to, w + off[i], esz[i] * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!stot) return MCS_OK;
    // staged columns as one byte range split over host threads
    parallel_rows((int64_t)stot, [&](int64_t a, int64_t b) {
        for (int i = 0; i < k; ++i) {
            if (!dst[i] || pinned[i]) continue;
            const int64_t s0 = (int64_t)soff[i], s1 = s0 + (int64_t)(esz[i] * n);
            const int64_t lo = std::max(a, s0), hi = std::min(b, s1);
            if (lo < hi) memcpy((char *)dst[i] + (lo - s0), (const char *)c->h_stage + lo, (size_t)(hi - lo));
        }
    });
    return MCS_OK;
}

extern "C" {

int mcs_create(int device, mcs_ctx **out)
{
    if (!out) return MCS_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MCS_ERR_HIP;
    if (device < 0 || device >= n) return MCS_ERR_ARG;
    mcs_ctx *c = new (std::nothrow) mcs_ctx();
    if (!c) return MCS_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return MCS_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return MCS_ERR_HIP; }
    c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void **)&c->h_sum, 8 * sizeof(unsigned long long), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&c->d_sum, c->h_sum, 0) != hipSuccess) {
        (void)mcs_destroy(c);
        return MCS_ERR_HIP;
    }
    // the no-alloc walk's scratch on this device (looked up once per device);
    // a failed lookup only disables mcs_walk_noalloc
    c->noalloc_err = noalloc_base(device, &c->noalloc_scratch);
    if (c->noalloc_err != hipSuccess) c->noalloc_scratch = nullptr;
    *out = c;
    return MCS_OK;
}

int mcs_destroy(mcs_ctx *c)
{
    if (!c) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    DevBuf *all[] = {&c->d_cases, &c->d_bands, &c->d_chunk0, &c->d_tally_f, &c->d_tally_u,
                     &c->d_ctl, &c->d_pband, &c->d_rdev, &c->d_roff, &c->d_pm, &c->d_pmf, &c->d_stokes, &c->d_wide};
    for (DevBuf *b : all) if (b->p) (void)hipFree(b->p);
    for (auto &b : c->d_rec) if (b.p) (void)hipFree(b.p);
    for (auto &b : c->d_out) if (b.p) (void)hipFree(b.p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->h_sum) (void)hipHostFree(c->h_sum);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MCS_OK;
}

const char *mcs_last_error(const mcs_ctx *c) { return c ? c->err : "null context"; }

int mcs_device_cus(const mcs_ctx *c, int32_t *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->cus;
    return MCS_OK;
}

int mcs_set_cases(mcs_ctx *c, const mcs_case *cases, int32_t n_cases, const mcs_band *bands,
                  int32_t n_bands)
{
    if (!c) return MCS_ERR_ARG;
    if (!cases || !bands || n_cases <= 0 || n_bands <= 0)
        return set_err(c, MCS_ERR_ARG, "cases/bands must be non-empty");
    std::vector<DevCase> dc(n_cases);
    std::vector<DevBand> db(n_bands);
    for (int b = 0; b < n_bands; ++b) {
        const mcs_band &B = bands[b];
        const double v[] = {B.wvn, B.g, B.ssa_ice, B.ssa_imp, B.ext_cff, B.p_ext_imp, B.snow_depth, B.cdf};
        for (double x : v)
            if (!std::isfinite(x)) return set_err(c, MCS_ERR_ARG, "band %d: non-finite value", b);
        if (!(B.g > -1.0 && B.g < 1.0)) return set_err(c, MCS_ERR_ARG, "band %d: |g| must be < 1", b);
        if (B.ssa_ice < 0 || B.ssa_ice > 1 || B.ssa_imp < 0 || B.ssa_imp > 1)
            return set_err(c, MCS_ERR_ARG, "band %d: ssa outside [0, 1]", b);
        if (!(B.ext_cff > 0)) return set_err(c, MCS_ERR_ARG, "band %d: ext_cff must be > 0", b);
        if (B.p_ext_imp < 0 || B.p_ext_imp > 1) return set_err(c, MCS_ERR_ARG, "band %d: p_ext_imp outside [0, 1]", b);
        if (B.wvn <= 0) return set_err(c, MCS_ERR_ARG, "band %d: wvn must be > 0", b);
        DevBand &D = db[b];
        D.wvn = B.wvn; D.g = B.g; D.ssa_ice = B.ssa_ice; D.ssa_imp = B.ssa_imp;
        D.ext_cff = B.ext_cff; D.p_ext_imp = B.p_ext_imp; D.cdf = B.cdf;
        D.snow_depth = B.snow_depth;
        D.iso = (B.g == 0.0);
        const double g = D.iso ? 0.5 : B.g;
        D.hg_a = (float)(1.0 / (2.0 * g)); D.hg_b = (float)(1.0 + g * g); D.hg_c = (float)(1.0 - g * g);
        D.hg_d = D.iso ? 1.f : (float)(1.0 - g + 2.0 * g * 0x1p-33);
        D.hg_e = D.iso ? 0.f : (float)(2.0 * g * 0x1p-32);
        D.inv_ext = (float)(1.0 / B.ext_cff);
        const double q5 = B.p_ext_imp * (1.0 - B.ssa_imp);
        D.t_abs5 = prob_bits(q5);
        D.t_abs_all = prob_bits(q5 + (1.0 - B.p_ext_imp) * (1.0 - B.ssa_ice));
        D.pm = -1;   // Henyey-Greenstein until mcs_set_polar
    }
    int64_t total = 0;
    std::vector<int> row_case(n_bands, -1);   // first case using each band row
    for (int k = 0; k < n_cases; ++k) {
        const mcs_case &C = cases[k];
        if (!(C.tau_tot > 0) || !std::isfinite(C.tau_tot)) return set_err(c, MCS_ERR_ARG, "case %d: tau_tot must be finite and > 0", k);
        if (!std::isfinite(C.mux0) || !std::isfinite(C.muz0) || !(C.muz0 < 0) ||
            std::fabs(C.mux0 * C.mux0 + C.muz0 * C.muz0 - 1.0) > 1e-9)
            return set_err(c, MCS_ERR_ARG, "case %d: (mux0, muz0) must be a downward unit vector", k);
        if (!(C.r_lambertian >= 0 && C.r_lambertian <= 1)) return set_err(c, MCS_ERR_ARG, "case %d: r_lambertian outside [0, 1]", k);
        if (C.n_bands < 1 || C.n_bands > MCS_MAX_BANDS || C.band_begin < 0 || C.band_begin + C.n_bands > n_bands)
            return set_err(c, MCS_ERR_ARG, "case %d: band range [%d, %d) invalid", k, C.band_begin, C.band_begin + C.n_bands);
        if (C.n_photons < 0 || C.n_photons > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "case %d: n_photons outside [0, 2^32)", k);
        // a band row's cdf (and alias entry) belongs to one band range
        for (int j = C.band_begin; j < C.band_begin + C.n_bands; ++j) {
            const int o = row_case[j];
            if (o < 0) { row_case[j] = k; continue; }
            if (cases[o].band_begin != C.band_begin || cases[o].n_bands != C.n_bands)
                return set_err(c, MCS_ERR_ARG, "cases %d and %d share band %d with different band ranges", o, k, j);
        }
        DevCase &D = dc[k];
        D.tau_tot = C.tau_tot; D.mux0 = C.mux0; D.muz0 = C.muz0; D.r_lamb = C.r_lambertian;
        D.t_refl = (unsigned long long)llround(C.r_lambertian * 4294967296.0);
        D.tau_tot_f = (float)C.tau_tot; D.mux0_f = (float)C.mux0; D.muz0_f = (float)C.muz0;
        D.lamb = C.lambertian_bottom ? 1 : 0; D.band_begin = C.band_begin; D.n_bands = C.n_bands;
        D.surf = C.lambertian_surface ? 1 : 0;
        D.polar = 0;
        D.s0_exact = 0;
        D.s0[0] = 1.0; D.s0f[0] = 1.f;   // (1, 0, 0, 0): unpolarised until mcs_set_polar
        D.n_photons = C.n_photons; D.gid_base = total;
        total += C.n_photons;
        build_alias(&db[C.band_begin], &bands[C.band_begin], C.n_bands);
    }
    (void)hipSetDevice(c->device);
    if (c->pending) (void)hipStreamSynchronize(c->stream), c->pending = false;
    // until the new tables are on the device the context holds no cases: a
    // failure below (e.g. MCS_ERR_NOMEM after a buffer was freed to grow it)
    // leaves it in that state, so a later walk reports MCS_ERR_STATE
    c->cases_set = false;
    c->polar = false;
    c->last_records = c->last_stokes = false;
    int r;
    if ((r = ensure(c, c->d_cases, sizeof(DevCase) * n_cases))) return r;
    if ((r = ensure(c, c->d_bands, sizeof(DevBand) * n_bands))) return r;
    if ((r = ensure(c, c->d_tally_f, sizeof(double) * TF * n_cases))) return r;
    if ((r = ensure(c, c->d_tally_u, sizeof(unsigned long long) * TU * n_cases))) return r;
    if ((r = ensure(c, c->d_chunk0, sizeof(int32_t) * (n_cases + 1)))) return r;
    c->chunk0_valid = false;
    if ((r = ensure(c, c->d_ctl, sizeof(unsigned long long) * 16))) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, dc.data(), sizeof(DevCase) * n_cases, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, db.data(), sizeof(DevBand) * n_bands, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cases.assign(cases, cases + n_cases);
    c->bands.assign(bands, bands + n_bands);
    c->dcases.swap(dc);
    c->dbands.swap(db);
    c->total_photons = total;
    c->cases_set = true;
    c->polar = false;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_set_polar(mcs_ctx *c, const mcs_polar_case *pc, int32_t n_cases, const double *tables,
                  int32_t n_tables, int32_t n_angles)
{
    if (!c) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    const int nc = (int)c->cases.size();
    std::vector<DevCase> dc = c->dcases;
    std::vector<DevBand> db = c->dbands;
    for (auto &D : dc) {
        D.polar = 0;
        D.s0_exact = 0;
        D.s0[0] = 1.0; D.s0[1] = D.s0[2] = D.s0[3] = 0.0;
        D.s0f[0] = 1.f; D.s0f[1] = D.s0f[2] = D.s0f[3] = 0.f;
    }
    for (auto &D : db) D.pm = -1;
    bool any = false;
    const int n = n_angles;
    if (pc) {
        if (n_cases != nc) return set_err(c, MCS_ERR_ARG, "n_cases = %d, %d cases are set", n_cases, nc);
        if (n_tables < 0 || (n_tables > 0 && (!tables || n < 2)))
            return set_err(c, MCS_ERR_ARG, "tables: need n_tables >= 0 and n_angles >= 2");
        if ((int64_t)n_tables * 7 * n > ((int64_t)1 << 31))
            return set_err(c, MCS_ERR_ARG, "phase-matrix tables too large");
        for (int t = 0; t < n_tables; ++t) {
            const double *x = tables + (size_t)t * 7 * n;
            for (int k = 0; k < 7 * n; ++k)
                if (!std::isfinite(x[k])) return set_err(c, MCS_ERR_ARG, "table %d: non-finite value", t);
            for (int k = 0; k < n; ++k) {
                if (x[k] < -1.0 || x[k] > 1.0) return set_err(c, MCS_ERR_ARG, "table %d: cos theta outside [-1, 1]", t);
                if (k && !(x[k] > x[k - 1])) return set_err(c, MCS_ERR_ARG, "table %d: cos theta not ascending", t);
            }
        }
        for (int k = 0; k < nc; ++k) {
            const mcs_polar_case &Q = pc[k];
            if (Q.table < 0) continue;
            DevCase &D = dc[k];
            if (Q.table + D.n_bands > n_tables)
                return set_err(c, MCS_ERR_ARG, "case %d: tables [%d, %d) beyond %d", k, Q.table, Q.table + D.n_bands, n_tables);
            if (Q.cutoff < 0 || Q.cutoff >= n) return set_err(c, MCS_ERR_ARG, "case %d: cutoff outside the grid", k);
            for (int q = 0; q < 4; ++q)
                if (!std::isfinite(Q.stokes0[q])) return set_err(c, MCS_ERR_ARG, "case %d: non-finite Stokes vector", k);
            if (!(Q.stokes0[0] != 0.0)) return set_err(c, MCS_ERR_ARG, "case %d: Stokes I must be non-zero", k);
            const double *x0 = tables + (size_t)Q.table * 7 * n;
            const int jc = n - 1 - Q.cutoff;   // the cutoff in ascending order
            D.polar = 1;
            // the exact sampler needs a physical initial vector (I > 0, degree
            // of polarisation <= 1); other cases take the two-area rejection
            const double s0i = Q.stokes0[0];
            D.s0_exact = (s0i > 0.0 && Q.stokes0[1] * Q.stokes0[1] + Q.stokes0[2] * Q.stokes0[2] +
                                           Q.stokes0[3] * Q.stokes0[3] <= s0i * s0i) ? 1 : 0;
            for (int q = 0; q < 4; ++q) { D.s0[q] = Q.stokes0[q]; D.s0f[q] = (float)Q.stokes0[q]; }
            D.cmax = x0[n - 1]; D.cmin = x0[0]; D.ccut = x0[jc];
            D.fcmax = (float)D.cmax; D.fcmin = (float)D.cmin; D.fccut = (float)D.ccut;
            for (int b = 0; b < D.n_bands; ++b) {
                const int t = Q.table + b;
                const double *x = tables + (size_t)t * 7 * n;
                for (int k2 = 0; k2 < n; ++k2)
                    if (x[k2] != x0[k2]) return set_err(c, MCS_ERR_ARG, "case %d: its tables have different grids", k);
                DevBand &B = db[D.band_begin + b];
                if (B.pm >= 0 && B.pm != t)
                    return set_err(c, MCS_ERR_ARG, "band row %d: two phase-matrix tables", D.band_begin + b);
                B.pm = t;
                // |P11|, |P12| maxima over theta <= theta_cut (area 1) and >= (area 2)
                double a1 = 0, b1 = 0, a2 = 0, b2 = 0;
                for (int k2 = 0; k2 < n; ++k2) {
                    const double p11 = std::fabs(x[n + k2]), p12 = std::fabs(x[2 * n + k2]);
                    if (k2 >= jc) { a1 = std::max(a1, p11); b1 = std::max(b1, p12); }
                    if (k2 <= jc) { a2 = std::max(a2, p11); b2 = std::max(b2, p12); }
                }
                if (!(a1 > 0) || !(a2 > 0)) return set_err(c, MCS_ERR_ARG, "table %d: P11 must be positive", t);
                B.pA1 = a1; B.pB1 = b1; B.pA2 = a2; B.pB2 = b2;
                B.fA1 = (float)a1; B.fB1 = (float)b1; B.fA2 = (float)a2; B.fB2 = (float)b2;
                bool exact = true;
                double mass = 0.0;
                for (int k2 = 0; k2 < n; ++k2) {
                    if (!(x[n + k2] >= 0.0) || std::fabs(x[2 * n + k2]) > x[n + k2] ||
                        !keeps_physical(x[n + k2], x[2 * n + k2], x[3 * n + k2], x[4 * n + k2], x[5 * n + k2],
                                        x[6 * n + k2]))
                        exact = false;
                    if (k2 + 1 < n) mass += 0.5 * (x[n + k2] + x[n + k2 + 1]) * (x[k2 + 1] - x[k2]);
                }
                B.pm_exact = (exact && mass > 0.0) ? 1 : 0;
                B.fM = (float)mass;
            }
            any = true;
        }
    }
    (void)hipSetDevice(c->device);
    if (c->pending) (void)hipStreamSynchronize(c->stream), c->pending = false;
    // a failure below may leave device tables that disagree with the host
    // copies (or freed): the context then holds no cases until mcs_set_cases
    c->cases_set = false;
    c->last_records = c->last_stokes = false;
    int r;
    if (any) {
        const size_t nv = (size_t)n_tables * 7 * n;
        const size_t stride = (size_t)14 * n + PM_GUIDE + 1;
        std::vector<float> tf((size_t)n_tables * stride, 0.f);
        std::vector<double> cum(n);
        for (int t = 0; t < n_tables; ++t) {
            const double *x = tables + (size_t)t * 7 * n;
            float *f = tf.data() + (size_t)t * stride;
            for (int k = 0; k < 7 * n; ++k) f[k] = (float)x[k];
            cum[0] = 0.0;   // integral of the linear P11 from x[0]
            for (int k = 0; k + 1 < n; ++k) cum[k + 1] = cum[k] + 0.5 * (x[n + k] + x[n + k + 1]) * (x[k + 1] - x[k]);
            for (int k = 0; k < n; ++k) f[7 * n + k] = (float)cum[k];
            for (int e = 0; e < 6; ++e)
                for (int k = 0; k + 1 < n; ++k)
                    f[(8 + e) * n + k] = (float)((x[(1 + e) * n + k + 1] - x[(1 + e) * n + k]) / (x[k + 1] - x[k]));
            int *guide = reinterpret_cast<int *>(f + 14 * n);
            int j = 0;
            for (int g = 0; g <= PM_GUIDE; ++g) {
                const double target = cum[n - 1] * g / PM_GUIDE;
                while (j < n - 2 && cum[j + 1] <= target) ++j;
                guide[g] = j;
            }
            guide[PM_GUIDE] = n - 2;
        }
        if ((r = ensure(c, c->d_pm, sizeof(double) * nv))) return r;
        if ((r = ensure(c, c->d_pmf, sizeof(float) * tf.size()))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_pm.p, tables, sizeof(double) * nv, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_pmf.p, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, dc.data(), sizeof(DevCase) * nc, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, db.data(), sizeof(DevBand) * db.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dcases.swap(dc);
    c->dbands.swap(db);
    c->cases_set = true;
    c->polar = any;
    c->pm_n = any ? n : 0;
    c->pm_floats = any ? (size_t)n_tables * ((size_t)14 * n + PM_GUIDE + 1) : 0;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_set_launch(mcs_ctx *c, const mcs_launch *l)
{
    if (!c || !l) return MCS_ERR_ARG;
    if (l->block < 64 || l->block > 1024 || l->block % 64 || l->blocks_per_cu < 1 || l->blocks_per_cu > 64 ||
        l->chunk < 64 || l->chunk > (1 << 24) || l->refill_min < 1 || l->refill_min > 64 || l->coop_live < 0 || l->coop_live > 32)
        return set_err(c, MCS_ERR_ARG, "launch shape out of range");
    c->launch = *l;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_get_launch(const mcs_ctx *c, mcs_launch *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->launch;
    return MCS_OK;
}

static int finish_walk(mcs_ctx *c);

// Photons per chunk of a walk: the launch's, raised to the block's threads.
// A block holds at most its tally slots' worth of chunks, each until its
// longest walk ends; chunks much smaller than the block left most of its
// lanes idle while the queue still had work (489-photon chunks in
// 1024-thread blocks: 6.3 ms for the 1e6-photon case, 0.7 ms at 977).
static int walk_chunk(const mcs_ctx *c) { return std::max(c->launch.chunk, c->launch.block); }

static uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

// enqueue fill + walk + scatter on the context stream
static int enqueue_walk(mcs_ctx *c, int mode, const KParams &P, uint32_t n_chunks, int64_t n, bool records,
                        bool events)
{
    const int nc = (int)c->cases.size();
    const uint64_t nclear = (uint64_t)(TF + TU) * nc + 16;
    const unsigned blocks_clear = (unsigned)std::min<uint64_t>((nclear + 255) / 256, 1024);
    hipLaunchKernelGGL(k_clear, dim3(blocks_clear), dim3(256), 0, c->stream,
                       (unsigned long long *)P.tally_f, (uint64_t)TF * nc, P.tally_u, (uint64_t)TU * nc,
                       P.ctl, (uint64_t)16);
    int block = (mode == MCS_REPLAY) ? std::min(c->launch.block, ReplayWalker::MAX_BLOCK)
                                     : c->launch.block;
    int grid = c->launch.blocks_per_cu * c->cus;
    if (mode == MCS_NATIVE && c->polar && block > PolarWalker::MAX_BLOCK) {
        grid *= block / PolarWalker::MAX_BLOCK;   // same threads per CU in smaller blocks
        block = PolarWalker::MAX_BLOCK;
    }
    if ((uint32_t)grid > n_chunks) grid = (int)std::max<uint32_t>(n_chunks, 1u);
    // events bracket the walk kernel alone: kernel_ms is its device time
    if (events && hipEventRecord(c->ev0, c->stream) != hipSuccess)
        return set_err(c, MCS_ERR_HIP, "hipEventRecord failed");
    if (mode == MCS_NATIVE && c->polar) {
        // aspherical particles: one photon per lane, then the tail mode of
        // polar_tail; tables up to 16 KB ride in LDS
        // + the tail mode's owner lanes and packets (16 ints and PK_SLOTS packets a wave)
        const size_t lds = (size_t)P.pm_lds_floats * sizeof(float) +
                           (size_t)(block / 64) * (16 * sizeof(int) + PK_SLOTS * sizeof(PolarPacket)) + 16;
        // blocks of up to 256 threads: two chunk slots (k_walk's SL)
        if (P.rec_direct) {
            if (block <= 256)
                hipLaunchKernelGGL((k_walk<PolarWalker, 3, true, 2>), dim3(grid), dim3(block), lds, c->stream, P);
            else
                hipLaunchKernelGGL((k_walk<PolarWalker, 3, true>), dim3(grid), dim3(block), lds, c->stream, P);
        } else {
            if (block <= 256)
                hipLaunchKernelGGL((k_walk<PolarWalker, 3, false, 2>), dim3(grid), dim3(block), lds, c->stream, P);
            else
                hipLaunchKernelGGL((k_walk<PolarWalker, 3>), dim3(grid), dim3(block), lds, c->stream, P);
        }
    } else if (mode == MCS_NATIVE) {
        // the instance whose register budget fits the shape's waves per SIMD:
        // <= 4 (1024-thread blocks, one per CU) 128 VGPRs, 5 96, more 80 (the
        // 6-wave build's spills sit outside the step loop)
        const int waves_per_simd = (c->launch.blocks_per_cu * (block / 64) + 3) / 4;
        // cooperative tables, or at coop_live 0 the packet tail's owners and
        // packets (in the build that carries it)
        const size_t tab = c->launch.coop_live > 0
                               ? (size_t)(block / 64) * c->launch.coop_live * (sizeof(CoopEntry) + sizeof(int))
                           : waves_per_simd <= 4
                               ? (size_t)(block / 64) * (HG_TAIL * sizeof(int) + HG_PK_SLOTS * sizeof(HgPacket))
                               : 0;
        if (waves_per_simd > 5 && P.rec_direct)
            hipLaunchKernelGGL((k_walk<NativeWalker, 6, true>), dim3(grid), dim3(block), tab, c->stream, P);
        else if (waves_per_simd > 5)
            hipLaunchKernelGGL((k_walk<NativeWalker, 6>), dim3(grid), dim3(block), tab, c->stream, P);
        else if (waves_per_simd <= 4 && P.rec_direct)
            hipLaunchKernelGGL((k_walk<NativeWalker, 4, true>), dim3(grid), dim3(block), tab, c->stream, P);
        else if (waves_per_simd <= 4)
            hipLaunchKernelGGL((k_walk<NativeWalker, 4>), dim3(grid), dim3(block), tab, c->stream, P);
        else if (P.rec_direct)
            hipLaunchKernelGGL((k_walk<NativeWalker, 5, true>), dim3(grid), dim3(block), tab, c->stream, P);
        else
            hipLaunchKernelGGL((k_walk<NativeWalker, 5>), dim3(grid), dim3(block), tab, c->stream, P);
    } else {
        hipLaunchKernelGGL((k_walk<ReplayWalker, ReplayWalker::WAVES_PER_EU>), dim3(grid), dim3(block), 0,
                           c->stream, P);
    }
    if (events && hipEventRecord(c->ev1, c->stream) != hipSuccess)
        return set_err(c, MCS_ERR_HIP, "hipEventRecord failed");
    const size_t rsz = (mode == MCS_NATIVE) ? sizeof(float) : sizeof(double);
    const size_t ord_lds = (size_t)P.chunk * (8 + 3 * rsz);
    if (records && n > 0 && ord_lds <= 65536) {
        const unsigned g2 = std::min<uint32_t>(n_chunks, 4096u);
        if (mode == MCS_NATIVE)
            hipLaunchKernelGGL(k_order_records<float>, dim3(g2), dim3(256), ord_lds, c->stream,
                               P.cases, P.case_chunk0, nc, (int)P.chunk, n_chunks,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (float *)c->d_rec[3].p, (float *)c->d_rec[4].p, (float *)c->d_rec[5].p,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (float *)c->d_out[2].p,
                               (float *)c->d_out[3].p, (float *)c->d_out[4].p);
        else
            hipLaunchKernelGGL(k_order_records<double>, dim3(g2), dim3(256), ord_lds, c->stream,
                               P.cases, P.case_chunk0, nc, (int)P.chunk, n_chunks,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (double *)c->d_rec[3].p, (double *)c->d_rec[4].p, (double *)c->d_rec[5].p,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (double *)c->d_out[2].p,
                               (double *)c->d_out[3].p, (double *)c->d_out[4].p);
    } else if (records && n > 0) {
        // chunks too large for LDS: scatter record by record
        const unsigned g2 = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
        if (mode == MCS_NATIVE)
            hipLaunchKernelGGL(k_scatter_records<float>, dim3(g2), dim3(256), 0, c->stream,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (float *)c->d_rec[3].p, (float *)c->d_rec[4].p, (float *)c->d_rec[5].p, (uint64_t)n,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (float *)c->d_out[2].p,
                               (float *)c->d_out[3].p, (float *)c->d_out[4].p);
        else
            hipLaunchKernelGGL(k_scatter_records<double>, dim3(g2), dim3(256), 0, c->stream,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (double *)c->d_rec[3].p, (double *)c->d_rec[4].p, (double *)c->d_rec[5].p, (uint64_t)n,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (double *)c->d_out[2].p,
                               (double *)c->d_out[3].p, (double *)c->d_out[4].p);
    }
    hipLaunchKernelGGL(k_summary, dim3(1), dim3(256), 0, c->stream, (const unsigned long long *)P.tally_u, nc,
                       (const unsigned long long *)P.ctl, c->d_sum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, MCS_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
    return MCS_OK;
}

int mcs_walk(mcs_ctx *c, const mcs_walk_args *a)
{
    if (!c || !a) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    if (a->mode != MCS_NATIVE && a->mode != MCS_REPLAY) return set_err(c, MCS_ERR_ARG, "unknown mode %d", a->mode);
    if (a->max_steps < 0 || a->max_steps > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "max_steps outside [0, 2^32)");
    (void)hipSetDevice(c->device);
    if (c->pending) {
        int r = finish_walk(c);
        if (r) return r;
    }
    const int64_t n = c->total_photons;
    const bool records = a->want_records != 0;
    if (records && n > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "records need fewer than 2^32 photons");
    if (a->mode == MCS_REPLAY) {
        if (!a->replay_dev || !a->replay_off || !a->photon_band)
            return set_err(c, MCS_ERR_ARG, "replay needs replay_dev, replay_off and photon_band");
    }
    const int nc = (int)c->cases.size();
    // photon bands: validate against each case's band count
    if (a->photon_band) {
        for (int k = 0; k < nc; ++k) {
            const DevCase &D = c->dcases[k];
            for (int64_t j = 0; j < D.n_photons; ++j)
                if (a->photon_band[D.gid_base + j] >= D.n_bands)
                    return set_err(c, MCS_ERR_ARG, "photon %lld: band %d >= n_bands %d of case %d",
                                   (long long)(D.gid_base + j), a->photon_band[D.gid_base + j], D.n_bands, k);
        }
    }
    int64_t rlen = 0;
    if (a->mode == MCS_REPLAY) {
        if (a->replay_off[0] != 0) return set_err(c, MCS_ERR_ARG, "replay_off[0] must be 0");
        for (int64_t j = 0; j < n; ++j)
            if (a->replay_off[j + 1] < a->replay_off[j] + 3)
                return set_err(c, MCS_ERR_ARG, "replay slice %lld shorter than the initial triple", (long long)j);
        rlen = a->replay_off[n];
    }
    // chunking
    const int chunk = walk_chunk(c);
    std::vector<int32_t> chunk0(nc + 1);
    int64_t nch = 0;
    for (int k = 0; k < nc; ++k) {
        chunk0[k] = (int32_t)nch;
        nch += (c->dcases[k].n_photons + chunk - 1) / chunk;
        if (nch > 0x7FFFFFFF) return set_err(c, MCS_ERR_ARG, "too many chunks; raise the chunk size");
    }
    chunk0[nc] = (int32_t)nch;
    // from here on the previous walk's records are no longer readable (their
    // buffers may be regrown below, and a failure leaves none)
    c->last_records = c->last_stokes = false;
    int r;
    if (!c->chunk0_valid || c->chunk0 != chunk0) {
        c->chunk0 = chunk0;
        c->chunk0_valid = false;
        HIPCHK(c, hipMemcpyAsync(c->d_chunk0.p, c->chunk0.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice,
                                 c->stream));
        c->chunk0_valid = true;
    }
    if (a->photon_band) {
        if ((r = ensure(c, c->d_pband, (size_t)n))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_pband.p, a->photon_band, (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    if (a->mode == MCS_REPLAY) {
        if ((r = ensure(c, c->d_rdev, sizeof(double) * (size_t)rlen))) return r;
        if ((r = ensure(c, c->d_roff, sizeof(int64_t) * (size_t)(n + 1)))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_rdev.p, a->replay_dev, sizeof(double) * (size_t)rlen, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_roff.p, a->replay_off, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
    }
    const size_t rsz = (a->mode == MCS_NATIVE) ? 4 : 8;
    if (records) {
        for (int k = 0; k < 3; ++k) if ((r = ensure(c, c->d_rec[k], 4 * (size_t)n))) return r;
        for (int k = 3; k < 6; ++k) if ((r = ensure(c, c->d_rec[k], rsz * (size_t)n))) return r;
        for (int k = 0; k < 2; ++k) if ((r = ensure(c, c->d_out[k], 4 * (size_t)n))) return r;
        for (int k = 2; k < 5; ++k) if ((r = ensure(c, c->d_out[k], rsz * (size_t)n))) return r;
    }
    KParams P;
    memset(&P, 0, sizeof P);
    P.cases = (const DevCase *)c->d_cases.p;
    P.bands = (const DevBand *)c->d_bands.p;
    P.case_chunk0 = (const int32_t *)c->d_chunk0.p;
    P.n_cases = nc;
    P.chunk = chunk;
    P.n_chunks = (uint32_t)nch;
    P.refill_min = (uint32_t)c->launch.refill_min;
    P.coop_live = c->launch.coop_live;
    P.key0 = (uint32_t)a->seed;
    P.key1 = (uint32_t)(a->seed >> 32);
    P.max_steps = a->max_steps ? (uint32_t)a->max_steps : (1u << 24);
    P.photon_band = a->photon_band ? (const uint8_t *)c->d_pband.p : nullptr;
    P.rdev = (const double *)c->d_rdev.p;
    P.roff = (const int64_t *)c->d_roff.p;
    P.tally_f = (double *)c->d_tally_f.p;
    P.tally_u = (unsigned long long *)c->d_tally_u.p;
    P.ctl = (unsigned long long *)c->d_ctl.p;
    P.chunk_counter = (unsigned int *)((unsigned long long *)c->d_ctl.p + 8);
    P.pm = (const double *)c->d_pm.p;
    P.pmf = (const float *)c->d_pmf.p;
    P.pm_n = c->pm_n;
    P.pm_stride = c->pm_n * 14 + PM_GUIDE + 1;
    P.pm_lds_floats = (c->polar && c->pm_floats * sizeof(float) <= 16 * 1024) ? (int32_t)c->pm_floats : 0;
    P.n_total = (uint64_t)n;
    if (records && c->polar) {
        if ((r = ensure(c, c->d_stokes, 3 * rsz * (size_t)std::max<int64_t>(n, 1)))) return r;
        P.rec_stokes = c->d_stokes.p;
    }
    if (records) {
        P.rec_gid = (uint32_t *)c->d_rec[0].p;
        P.rec_nscat = (uint32_t *)c->d_rec[1].p;
        P.rec_cb = (uint32_t *)c->d_rec[2].p;
        P.rec_theta = c->d_rec[3].p;
        P.rec_phi = c->d_rec[4].p;
        P.rec_path = c->d_rec[5].p;
    }
    c->h_sum[7] = 0ull;   // set again by this walk's k_summary
    if (a->use_graph) {
        uint64_t key = 0x1234;
        const uint64_t *w = (const uint64_t *)&P;
        for (size_t k = 0; k < sizeof(P) / 8; ++k) key = mix64(key, w[k]);
        key = mix64(key, (uint64_t)a->mode);
        key = mix64(key, (uint64_t)records);
        key = mix64(key, (uint64_t)n);
        key = mix64(key, (uint64_t)c->launch.block * 131 + c->launch.blocks_per_cu);
        if (!c->gexec || c->gkey != key) {
            if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int er = enqueue_walk(c, a->mode, P, (uint32_t)nch, n, records, false);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(c->stream, &g);
            if (er) { if (g) (void)hipGraphDestroy(g); return er; }
            if (ee != hipSuccess) return set_err(c, MCS_ERR_HIP, "graph capture: %s", hipGetErrorString(ee));
            c->graph = g;
            HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
            c->gkey = key;
        }
        // a graph launch is timed as a whole (fills + walk + record scatter)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    } else {
        if ((r = enqueue_walk(c, a->mode, P, (uint32_t)nch, n, records, true))) return r;
    }
    c->last_mode = a->mode;
    c->last_records = records;
    c->last_stokes = records && c->polar;
    c->last_tallies_ctx = true;
    c->last_n = n;
    c->pending = true;
    if (a->async) return MCS_OK;
    return finish_walk(c);
}

int mcs_walk_noalloc(mcs_ctx *c, const mcs_walk_args *a, const mcs_dev_out *o)
{
    if (!c || !a || !o) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    if (a->mode != MCS_NATIVE) return set_err(c, MCS_ERR_ARG, "mcs_walk_noalloc walks in native mode only");
    if (a->photon_band || a->replay_dev || a->replay_off || a->use_graph)
        return set_err(c, MCS_ERR_ARG, "mcs_walk_noalloc takes no photon_band, replay input or graph launch");
    if (a->max_steps < 0 || a->max_steps > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "max_steps outside [0, 2^32)");
    const int nc = (int)c->cases.size();
    if (nc > MCS_NOALLOC_MAX_CASES)
        return set_err(c, MCS_ERR_ARG, "mcs_walk_noalloc: %d cases, at most %d (MCS_NOALLOC_MAX_CASES)", nc,
                       MCS_NOALLOC_MAX_CASES);
    if (!o->tallies || !o->counts) return set_err(c, MCS_ERR_ARG, "tallies and counts buffers are required");
    if (!c->noalloc_scratch)
        return set_err(c, MCS_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(c->noalloc_err));
    const int64_t n = c->total_photons;
    const bool records = a->want_records != 0;
    if (records) {
        if (!o->n_scat || !o->cond_band || !o->theta_n || !o->phi_n || !o->path_length)
            return set_err(c, MCS_ERR_ARG, "want_records needs all five record arrays");
        if (n > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "records need fewer than 2^32 photons");
    }
    (void)hipSetDevice(c->device);
    if (c->pending) {
        int r = finish_walk(c);
        if (r) return r;
    }
    const int chunk = walk_chunk(c);
    int64_t nch = 0;
    for (int k = 0; k < nc; ++k) nch += (c->dcases[k].n_photons + chunk - 1) / chunk;
    if (nch > 0x7FFFFFFF) return set_err(c, MCS_ERR_ARG, "too many chunks; raise the chunk size");
    c->last_records = c->last_stokes = false;
    unsigned long long *const scr = c->noalloc_scratch;   // this device's scratch
    int32_t *const chunk0 = reinterpret_cast<int32_t *>(scr + NOALLOC_CTL_U64);
    KParams P;
    memset(&P, 0, sizeof P);
    P.cases = (const DevCase *)c->d_cases.p;
    P.bands = (const DevBand *)c->d_bands.p;
    P.case_chunk0 = chunk0;
    P.n_cases = nc;
    P.chunk = chunk;
    P.n_chunks = (uint32_t)nch;
    P.refill_min = (uint32_t)c->launch.refill_min;
    P.coop_live = c->launch.coop_live;
    P.key0 = (uint32_t)a->seed;
    P.key1 = (uint32_t)(a->seed >> 32);
    P.max_steps = a->max_steps ? (uint32_t)a->max_steps : (1u << 24);
    P.tally_f = o->tallies;
    P.tally_u = (unsigned long long *)o->counts;
    P.ctl = scr;
    P.chunk_counter = (unsigned int *)(scr + 8);
    P.pm = (const double *)c->d_pm.p;
    P.pmf = (const float *)c->d_pmf.p;
    P.pm_n = c->pm_n;
    P.pm_stride = c->pm_n * 14 + PM_GUIDE + 1;
    P.pm_lds_floats = (c->polar && c->pm_floats * sizeof(float) <= 16 * 1024) ? (int32_t)c->pm_floats : 0;
    P.n_total = (uint64_t)n;
    if (records) {
        P.rec_direct = 1;
        P.rec_nscat = o->n_scat;
        P.rec_cb = o->cond_band;
        P.rec_theta = o->theta_n;
        P.rec_phi = o->phi_n;
        P.rec_path = o->path_length;
        if (c->polar) P.rec_stokes = o->stokes;
    }
    c->h_sum[7] = 0ull;   // set again by this walk's k_summary
    hipLaunchKernelGGL(k_chunk_table, dim3(1), dim3(1024), 0, c->stream, P.cases, nc, chunk, chunk0);
    int r = enqueue_walk(c, MCS_NATIVE, P, (uint32_t)nch, n, false, true);
    if (r) return r;
    c->last_mode = MCS_NATIVE;
    c->last_tallies_ctx = false;
    c->last_n = n;
    c->pending = true;
    if (a->async) return MCS_OK;
    return finish_walk(c);
}

static int finish_walk(mcs_ctx *c)
{
    c->pending = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    // k_summary wrote the sums into pinned host memory before the stream drained
    const volatile unsigned long long *h = c->h_sum;
    if (h[7] != 1ull) return set_err(c, MCS_ERR_HIP, "internal: walk summary not written");
    unsigned long long ctl[3] = {h[4], h[5], h[6]};
    mcs_stats s{};
    s.photons = h[0];
    s.steps = h[1];
    s.capped = h[2];
    s.phi_unbound = h[3];
    s.kernel_ms = ms;
    c->stats = s;
    if (ctl[1])
        return set_err(c, MCS_ERR_HIP, "internal: %llu failed photon hand-overs between waves",
                       (unsigned long long)ctl[1]);
    if (ctl[2])
        return set_err(c, MCS_ERR_ARG, "%llu scattering angles not accepted in %u phase-matrix rejection tries "
                       "(P11 + P12 terms not positive?)", (unsigned long long)ctl[2], POLAR_MAX_TRIES);
    if (s.photons != (uint64_t)c->last_n)
        return set_err(c, MCS_ERR_HIP, "walk finished %llu of %lld photons",
                       (unsigned long long)s.photons, (long long)c->last_n);
    if (c->last_mode == MCS_REPLAY && ctl[0])
        return set_err(c, MCS_ERR_REPLAY, "%llu photons did not consume exactly their replay slice",
                       (unsigned long long)ctl[0]);
    if (s.capped) return set_err(c, MCS_ERR_STEP_CAP, "%llu photons reached max_steps", (unsigned long long)s.capped);
    if (c->last_mode == MCS_REPLAY && s.phi_unbound)
        return set_err(c, MCS_ERR_PHI_UNBOUND,
                       "%llu photons end on an axis: the reference leaves phi_n unbound there",
                       (unsigned long long)s.phi_unbound);
    return MCS_OK;
}

int mcs_sync(mcs_ctx *c)
{
    if (!c) return MCS_ERR_ARG;
    if (!c->pending) return MCS_OK;
    (void)hipSetDevice(c->device);
    return finish_walk(c);
}

int mcs_read_records(mcs_ctx *c, const mcs_records *o, int64_t n)
{
    if (!c || !o) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_records) return set_err(c, MCS_ERR_STATE, "last walk kept no records");
    if (n != c->last_n) return set_err(c, MCS_ERR_ARG, "n = %lld, last walk had %lld photons", (long long)n, (long long)c->last_n);
    if (n == 0) return MCS_OK;
    (void)hipSetDevice(c->device);
    // the device writes every requested column in its host layout, then each
    // column is one copy: straight into the caller's array when that is
    // pinned host memory, else through pinned staging
    void *dst[8] = {o->condition, o->band, o->n_scat, o->theta_n, o->phi_n, o->path_length, o->wvn, o->snow_depth};
    const size_t esz[8] = {1, 1, 8, 8, 8, 8, 8, 8};
    size_t off[8] = {}, total = 0;
    for (int k = 0; k < 8; ++k)
        if (dst[k]) { off[k] = total; total += (esz[k] * (size_t)n + 255) & ~(size_t)255; }
    if (!total) return MCS_OK;
    int r;
    if ((r = ensure(c, c->d_wide, total))) return r;
    char *w = (char *)c->d_wide.p;
    WideRecords W;
    W.condition = dst[0] ? (int8_t *)(w + off[0]) : nullptr;
    W.band = dst[1] ? (uint8_t *)(w + off[1]) : nullptr;
    W.n_scat = dst[2] ? (int64_t *)(w + off[2]) : nullptr;
    W.theta_n = dst[3] ? (double *)(w + off[3]) : nullptr;
    W.phi_n = dst[4] ? (double *)(w + off[4]) : nullptr;
    W.path_length = dst[5] ? (double *)(w + off[5]) : nullptr;
    W.wvn = dst[6] ? (double *)(w + off[6]) : nullptr;
    W.snow_depth = dst[7] ? (double *)(w + off[7]) : nullptr;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 2048);
    const int nc = (int)c->cases.size();
    if (c->last_mode == MCS_REPLAY)
        hipLaunchKernelGGL(k_widen_records<double>, dim3(g), dim3(256), 0, c->stream, (const DevCase *)c->d_cases.p,
                           (const DevBand *)c->d_bands.p, nc, (uint64_t)n, (const uint32_t *)c->d_out[0].p,
                           (const uint32_t *)c->d_out[1].p, (const double *)c->d_out[2].p,
                           (const double *)c->d_out[3].p, (const double *)c->d_out[4].p, W);
    else
        hipLaunchKernelGGL(k_widen_records<float>, dim3(g), dim3(256), 0, c->stream, (const DevCase *)c->d_cases.p,
                           (const DevBand *)c->d_bands.p, nc, (uint64_t)n, (const uint32_t *)c->d_out[0].p,
                           (const uint32_t *)c->d_out[1].p, (const float *)c->d_out[2].p,
                           (const float *)c->d_out[3].p, (const float *)c->d_out[4].p, W);
    HIPCHK(c, hipGetLastError());
    return copy_columns(c, dst, esz, off, 8, (size_t)n);
}

int mcs_read_stokes(mcs_ctx *c, double *q, double *u, double *v, int64_t n)
{
    if (!c) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_stokes) return set_err(c, MCS_ERR_STATE, "last walk kept no Stokes vectors (records off or no polar cases)");
    if (n != c->last_n) return set_err(c, MCS_ERR_ARG, "n = %lld, last walk had %lld photons", (long long)n, (long long)c->last_n);
    if (n == 0) return MCS_OK;
    (void)hipSetDevice(c->device);
    int r;
    const size_t plane = (8 * (size_t)n + 255) & ~(size_t)255;
    if ((r = ensure(c, c->d_wide, 3 * plane))) return r;
    // the float64 planes are written plane-padded: widen each plane
    for (int k = 0; k < 3; ++k) {
        const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 2048);
        double *o = (double *)((char *)c->d_wide.p + k * plane);
        if (c->last_mode == MCS_REPLAY)
            hipLaunchKernelGGL(k_widen_stokes<double>, dim3(g), dim3(256), 0, c->stream,
                               (const double *)c->d_stokes.p + (size_t)k * n, (uint64_t)n, o);
        else
            hipLaunchKernelGGL(k_widen_stokes<float>, dim3(g), dim3(256), 0, c->stream,
                               (const float *)c->d_stokes.p + (size_t)k * n, (uint64_t)n, o);
    }
    HIPCHK(c, hipGetLastError());
    void *dst[3] = {q, u, v};
    const size_t esz[3] = {8, 8, 8};
    const size_t off[3] = {0, plane, 2 * plane};
    return copy_columns(c, dst, esz, off, 3, (size_t)n);
}

int mcs_read_tallies(mcs_ctx *c, double *out, int32_t n_cases)
{
    if (!c || !out) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_tallies_ctx) return set_err(c, MCS_ERR_STATE, "the last walk wrote its tallies to the caller's buffers");
    if (n_cases != (int32_t)c->cases.size()) return set_err(c, MCS_ERR_ARG, "n_cases mismatch");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_tally_f.p, sizeof(double) * TF * n_cases, hipMemcpyDeviceToHost));
    return MCS_OK;
}

int mcs_read_counts(mcs_ctx *c, uint64_t *out, int32_t n_cases)
{
    if (!c || !out) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_tallies_ctx) return set_err(c, MCS_ERR_STATE, "the last walk wrote its counts to the caller's buffers");
    if (n_cases != (int32_t)c->cases.size()) return set_err(c, MCS_ERR_ARG, "n_cases mismatch");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_tally_u.p, sizeof(uint64_t) * TU * n_cases, hipMemcpyDeviceToHost));
    return MCS_OK;
}

int mcs_last_stats(const mcs_ctx *c, mcs_stats *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->stats;
    return MCS_OK;
}

// Put back the phase-matrix state of the cases (mcs_set_cases clears it):
// the polar fields of the case and band rows, re-uploaded; the device tables
// themselves are untouched by mcs_set_cases.
static int restore_polar(mcs_ctx *c, const std::vector<DevCase> &dc, const std::vector<DevBand> &db, bool polar,
                         int32_t pm_n, size_t pm_floats)
{
    if (!polar) return MCS_OK;
    for (size_t k = 0; k < c->dcases.size(); ++k) {
        DevCase &D = c->dcases[k];
        const DevCase &S = dc[k];
        D.polar = S.polar;
        D.s0_exact = S.s0_exact;
        for (int q = 0; q < 4; ++q) { D.s0[q] = S.s0[q]; D.s0f[q] = S.s0f[q]; }
        D.cmax = S.cmax; D.ccut = S.ccut; D.cmin = S.cmin;
        D.fcmax = S.fcmax; D.fccut = S.fccut; D.fcmin = S.fcmin;
    }
    for (size_t k = 0; k < c->dbands.size(); ++k) {
        DevBand &B = c->dbands[k];
        const DevBand &S = db[k];
        B.pm = S.pm; B.pm_exact = S.pm_exact; B.fM = S.fM;
        B.pA1 = S.pA1; B.pB1 = S.pB1; B.pA2 = S.pA2; B.pB2 = S.pB2;
        B.fA1 = S.fA1; B.fB1 = S.fB1; B.fA2 = S.fA2; B.fB2 = S.fB2;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, c->dcases.data(), sizeof(DevCase) * c->dcases.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, c->dbands.data(), sizeof(DevBand) * c->dbands.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->polar = true;
    c->pm_n = pm_n;
    c->pm_floats = pm_floats;
    return MCS_OK;
}

int mcs_autotune(mcs_ctx *c, const mcs_launch *cand, int32_t n, int64_t photons_cap, int32_t repeats,
                 uint64_t seed, double *times_ms, mcs_launch *best)
{
    if (!c || !cand || n <= 0 || repeats <= 0 || !best) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    // a capped copy of the current cases
    std::vector<mcs_case> saved = c->cases;
    std::vector<mcs_band> bands = c->bands;
    std::vector<mcs_case> capped = saved;
    if (photons_cap > 0)
        for (auto &k : capped) k.n_photons = std::min<int64_t>(k.n_photons, photons_cap);
    const mcs_launch keep = c->launch;
    // the phase-matrix state survives the capped copy (aspherical cases are
    // timed with their own kernel)
    const std::vector<DevCase> pdc = c->dcases;
    const std::vector<DevBand> pdb = c->dbands;
    const bool polar = c->polar;
    const int32_t pm_n = c->pm_n;
    const size_t pm_floats = c->pm_floats;
    int r = mcs_set_cases(c, capped.data(), (int32_t)capped.size(), bands.data(), (int32_t)bands.size());
    if (r) return r;
    if ((r = restore_polar(c, pdc, pdb, polar, pm_n, pm_floats))) return r;
    double best_t = 1e300;
    int best_i = -1;
    for (int k = 0; k < n; ++k) {
        if ((r = mcs_set_launch(c, &cand[k]))) break;
        mcs_walk_args a;
        memset(&a, 0, sizeof a);
        a.mode = MCS_NATIVE;
        a.seed = seed;
        if ((r = mcs_walk(c, &a))) break;  // warm-up
        // the same seeds for every candidate, a different walk for every repeat
        std::vector<double> t;
        for (int q = 0; q < repeats; ++q) {
            a.seed = seed + 1 + (uint64_t)q;
            if ((r = mcs_walk(c, &a))) break;
            t.push_back(c->stats.kernel_ms);
        }
        if (r) break;
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (times_ms) times_ms[k] = med;
        if (med < best_t) { best_t = med; best_i = k; }
    }
    int r2 = mcs_set_cases(c, saved.data(), (int32_t)saved.size(), bands.data(), (int32_t)bands.size());
    if (!r2) r2 = restore_polar(c, pdc, pdb, polar, pm_n, pm_floats);
    if (r) { c->launch = keep; return r; }
    if (r2) return r2;
    *best = cand[best_i];
    c->launch = cand[best_i];
    return MCS_OK;
}

// Diagnostic: Philox4x32-10 of n counters (4 x u32 each) with key (k0, k1),
// computed on the device (checks the device generator against a host one).
int mcs_philox(mcs_ctx *c, const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out, int64_t n)
{
    if (!c || !ctr || !out || n <= 0) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    void *dc = nullptr, *dout = nullptr;
    HIPCHK(c, hipMalloc(&dc, 16 * (size_t)n));
    hipError_t e = hipMalloc(&dout, 16 * (size_t)n);
    if (e != hipSuccess) { (void)hipFree(dc); return set_err(c, MCS_ERR_NOMEM, "hipMalloc"); }
    e = hipMemcpy(dc, ctr, 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_philox, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           (const uint32_t *)dc, k0, k1, (uint32_t *)dout, (uint64_t)n);
        e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, 16 * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dout);
    if (e != hipSuccess) return set_err(c, MCS_ERR_HIP, "mcs_philox: %s", hipGetErrorString(e));
    return MCS_OK;
}

}  // extern "C"
